// Learned coarse correction: a 4-level 3D U-Net (BF16 weights, FP32 activations and
// accumulation) on the level-1 node grid, with hand-written forward, backward and Adam.
// No counterpart in the reference; the nearest pattern is its NumPy MLP with manual
// backward and Adam (deep_frame/topology_neural.py:59-113).
//
// Conventions (identical to the CPU restatement used by the tests):
//   activations a[n][c][X][Y][Z]; a 3x3x3 convolution with stride S reads input node
//   S*v + t - 1 per axis (zero outside), so stride 2 maps n nodes to (n - 1) / 2 + 1;
//   an up layer is the adjoint of a stride-2 convolution with weights [c_coarse][c_fine][27];
//   weights [c_out][c_in][27] are rounded to BF16 (nearest even) from an FP32 master copy.
#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>

#include "dfmi_internal.h"
#include "dfmi_learned.h"

namespace dfmi {

__device__ inline float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline uint16_t bf16_round(float f) {
  uint32_t b = __float_as_uint(f);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

enum Epilogue : int { EPI_NONE = 0, EPI_RELU = 1, EPI_MASK = 2 };

constexpr int CT = 16;   // output channels per thread
constexpr int CI = 16;   // input channels staged per LDS pass

__device__ inline void decode(long long v, const NetDims& d, int& x, int& y, int& z) {
  z = (int)(v % d.z);
  y = (int)((v / d.z) % d.y);
  x = (int)(v / ((long long)d.z * d.y));
}

// out[n][o][v] = sum_i sum_t W[o][i][t] in[n][i][S v + t - 1], then bias, add, epilogue
template <int S>
__global__ __launch_bounds__(256) void k_conv(const float* __restrict__ in, NetDims din, int ci,
                                              const uint16_t* __restrict__ w, int co, const float* bias,
                                              const float* add, const float* act, int epi, float* __restrict__ out,
                                              NetDims dout) {
  __shared__ float wl[CT][CI][27];
  const int groups = (co + CT - 1) / CT;
  const int n = blockIdx.y / groups, o0 = (blockIdx.y % groups) * CT;
  const long long nv = dout.n();
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool live = v < nv;
  int off[27];
  {
    int x = 0, y = 0, z = 0;
    if (live) decode(v, dout, x, y, z);
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int ix = S * x + t / 9 - 1, iy = S * y + (t / 3) % 3 - 1, iz = S * z + t % 3 - 1;
      const bool ok = live && ix >= 0 && ix < din.x && iy >= 0 && iy < din.y && iz >= 0 && iz < din.z;
      off[t] = ok ? (ix * din.y + iy) * din.z + iz : -1;
    }
  }
  float acc[CT];
#pragma unroll
  for (int q = 0; q < CT; ++q) acc[q] = 0.f;
  const long long plane = din.n();
  for (int i0 = 0; i0 < ci; i0 += CI) {
    for (int idx = threadIdx.x; idx < CT * CI * 27; idx += 256) {
      const int q = idx / (CI * 27), ii = (idx / 27) % CI, t = idx % 27;
      const int o = o0 + q, i = i0 + ii;
      wl[q][ii][t] = (o < co && i < ci) ? bf16_value(w[((long long)o * ci + i) * 27 + t]) : 0.f;
    }
    __syncthreads();
    const int nc = min(CI, ci - i0);
    for (int ii = 0; ii < nc; ++ii) {
      const float* p = in + ((long long)n * ci + i0 + ii) * plane;
#pragma unroll
      for (int t = 0; t < 27; ++t) {
        const float a = off[t] >= 0 ? p[off[t]] : 0.f;
#pragma unroll
        for (int q = 0; q < CT; ++q) acc[q] += wl[q][ii][t] * a;
      }
    }
    __syncthreads();
  }
  if (!live) return;
#pragma unroll
  for (int q = 0; q < CT; ++q) {
    const int o = o0 + q;
    if (o >= co) break;
    const long long idx = ((long long)n * co + o) * nv + v;
    float r = acc[q] + (bias ? bias[o] : 0.f) + (add ? add[idx] : 0.f);
    if (epi == EPI_RELU) r = fmaxf(r, 0.f);
    if (epi == EPI_MASK) r = act[idx] > 0.f ? r : 0.f;
    out[idx] = r;
  }
}

// Adjoint: out[n][i][u] = sum_o sum_t W[o][i][t] in[n][o][v] over S v + t - 1 = u (W is [no][ni][27]),
// then bias, add, epilogue.
template <int S>
__global__ __launch_bounds__(256) void k_conv_adj(const float* __restrict__ in, NetDims din, int no,
                                                  const uint16_t* __restrict__ w, int ni, const float* bias,
                                                  const float* add, const float* act, int epi,
                                                  float* __restrict__ out, NetDims dout) {
  __shared__ float wl[CT][CI][27];
  const int groups = (ni + CT - 1) / CT;
  const int n = blockIdx.y / groups, i0 = (blockIdx.y % groups) * CT;
  const long long nu = dout.n();
  const long long u = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool live = u < nu;
  int off[27];
  {
    int x = 0, y = 0, z = 0;
    if (live) decode(u, dout, x, y, z);
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int ax = x + 1 - t / 9, ay = y + 1 - (t / 3) % 3, az = z + 1 - t % 3;
      bool ok = live && ax >= 0 && ay >= 0 && az >= 0 && ax % S == 0 && ay % S == 0 && az % S == 0;
      const int vx = ax / S, vy = ay / S, vz = az / S;
      ok = ok && vx < din.x && vy < din.y && vz < din.z;
      off[t] = ok ? (vx * din.y + vy) * din.z + vz : -1;
    }
  }
  float acc[CT];
#pragma unroll
  for (int q = 0; q < CT; ++q) acc[q] = 0.f;
  const long long plane = din.n();
  for (int o0 = 0; o0 < no; o0 += CI) {
    for (int idx = threadIdx.x; idx < CT * CI * 27; idx += 256) {
      const int q = idx / (CI * 27), oo = (idx / 27) % CI, t = idx % 27;
      const int i = i0 + q, o = o0 + oo;
      wl[q][oo][t] = (o < no && i < ni) ? bf16_value(w[((long long)o * ni + i) * 27 + t]) : 0.f;
    }
    __syncthreads();
    const int nc = min(CI, no - o0);
    for (int oo = 0; oo < nc; ++oo) {
      const float* p = in + ((long long)n * no + o0 + oo) * plane;
#pragma unroll
      for (int t = 0; t < 27; ++t) {
        if (off[t] < 0) continue;
        const float a = p[off[t]];
#pragma unroll
        for (int q = 0; q < CT; ++q) acc[q] += wl[q][oo][t] * a;
      }
    }
    __syncthreads();
  }
  if (!live) return;
#pragma unroll
  for (int q = 0; q < CT; ++q) {
    const int i = i0 + q;
    if (i >= ni) break;
    const long long idx = ((long long)n * ni + i) * nu + u;
    float r = acc[q] + (bias ? bias[i] : 0.f) + (add ? add[idx] : 0.f);
    if (epi == EPI_RELU) r = fmaxf(r, 0.f);
    if (epi == EPI_MASK) r = act[idx] > 0.f ? r : 0.f;
    out[idx] = r;
  }
}

// partial[c][o][i][t] = sum over the c-th chunk of (n, v) of dy[n][o][v] x[n][i][S v + t - 1]
constexpr int WG_O = 8, WG_I = 8, WG_J = 32;
template <int S>
__global__ __launch_bounds__(WG_I * 27) void k_wgrad(const float* __restrict__ x, NetDims dx, int ni,
                                                     const float* __restrict__ dy, NetDims dyd, int no, int nsamples,
                                                     long long per_chunk, float* __restrict__ partial) {
  __shared__ float dl[WG_O][WG_J];
  const int c = blockIdx.x, o0 = blockIdx.y * WG_O, i0 = blockIdx.z * WG_I;
  const int il = threadIdx.x / 27, t = threadIdx.x % 27;
  const int i = i0 + il;
  const int tx = t / 9 - 1, ty = (t / 3) % 3 - 1, tz = t % 3 - 1;
  const long long nv = dyd.n(), total = (long long)nsamples * nv;
  const long long m_begin = (long long)c * per_chunk, m_end = min(total, m_begin + per_chunk);
  float acc[WG_O];
#pragma unroll
  for (int q = 0; q < WG_O; ++q) acc[q] = 0.f;
  const long long xplane = dx.n();
  for (long long m0 = m_begin; m0 < m_end; m0 += WG_J) {
    for (int idx = threadIdx.x; idx < WG_O * WG_J; idx += WG_I * 27) {
      const int q = idx / WG_J, j = idx % WG_J;
      const long long m = m0 + j;
      const int o = o0 + q;
      float val = 0.f;
      if (m < m_end && o < no) {
        const long long n = m / nv, v = m % nv;
        val = dy[(n * no + o) * nv + v];
      }
      dl[q][j] = val;
    }
    __syncthreads();
    const int cnt = (int)min((long long)WG_J, m_end - m0);
    for (int j = 0; j < cnt; ++j) {
      const long long m = m0 + j;
      const long long n = m / nv, v = m % nv;
      int vx, vy, vz;
      decode(v, dyd, vx, vy, vz);
      const int ix = S * vx + tx, iy = S * vy + ty, iz = S * vz + tz;
      float a = 0.f;
      if (i < ni && ix >= 0 && ix < dx.x && iy >= 0 && iy < dx.y && iz >= 0 && iz < dx.z)
        a = x[(n * ni + i) * xplane + ((long long)ix * dx.y + iy) * dx.z + iz];
#pragma unroll
      for (int q = 0; q < WG_O; ++q) acc[q] += dl[q][j] * a;
    }
    __syncthreads();
  }
  if (i >= ni) return;
#pragma unroll
  for (int q = 0; q < WG_O; ++q) {
    const int o = o0 + q;
    if (o < no) partial[(((long long)c * no + o) * ni + i) * 27 + t] = acc[q];
  }
}

// out[e] = sum_c partial[c][e] (fixed order)
__global__ void k_sum_chunks(const float* __restrict__ partial, int chunks, long long n, float* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.f;
  for (int c = 0; c < chunks; ++c) s += partial[(long long)c * n + e];
  out[e] = s;
}

// bias gradient: out[o] = sum_n sum_v dy[n][o][v]
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ dy, long long nv, int no, int nsamples,
                                                   float* __restrict__ out) {
  const int o = blockIdx.x;
  double s = 0.0;
  for (int n = 0; n < nsamples; ++n)
    for (long long v = threadIdx.x; v < nv; v += 256) s += dy[((long long)n * no + o) * nv + v];
  s = block_sum(s);
  if (threadIdx.x == 0) out[o] = (float)s;
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = bf16_round(w[i]);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wb, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  wb[i] = bf16_round(wi);
}

// ---- features, correction and loss on multigrid vectors (layout VLay, nodes of level 1) ----------
// per-sample max |Dinv^(1/2) g|
__global__ __launch_bounds__(256) void k_sample_absmax(const float* __restrict__ g, VLay lay, long long nodes,
                                                       const float* __restrict__ dinv, double* partial) {
  const int n = blockIdx.y;
  double m = 0.0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nodes * 3; e += (long long)gridDim.x * 256) {
    const long long node = e / 3;
    const int c = (int)(e % 3);
    m = fmax(m, (double)fabsf(g[node * lay.ns + n * lay.rs + c] * sqrtf(dinv[e])));
  }
  m = block_max(m);
  if (threadIdx.x == 0) partial[(long long)n * gridDim.x + blockIdx.x] = m;
}

// sum over nodes of a * (w ? w : 1) * b per sample
__global__ __launch_bounds__(256) void k_sample_dot(const float* __restrict__ a, const float* __restrict__ b,
                                                    const float* __restrict__ w, VLay lay, long long nodes,
                                                    double* partial) {
  const int n = blockIdx.y;
  double s = 0.0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nodes * 3; e += (long long)gridDim.x * 256) {
    const long long node = e / 3;
    const int c = (int)(e % 3);
    const long long o = node * lay.ns + n * lay.rs + c;
    s += (double)a[o] * (double)b[o] * (w ? (double)w[e] : 1.0);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[(long long)n * gridDim.x + blockIdx.x] = s;
}

// stiffness feature f[c][node] = log10(diag / diag_max) / 9 + 1 (0 on dead DOFs), diag = 1 / dinv
__global__ void k_stiffness_feature(const float* __restrict__ dinv, long long nodes, const double* dmax,
                                    float* __restrict__ f) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nodes * 3) return;
  const float di = dinv[e];
  const long long node = e / 3;
  const int c = (int)(e % 3);
  f[c * nodes + node] = di > 0.f ? (float)(log10(1.0 / ((double)di * dmax[0])) / 9.0 + 1.0) : 0.f;
}

__global__ void k_min_positive_partial(const float* __restrict__ dinv, long long n, double* partial) {
  // max of 1/dinv over positive entries, as a max-reduction of 1/dinv
  double m = 0.0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
    if (dinv[e] > 0.f) m = fmax(m, 1.0 / (double)dinv[e]);
  m = block_max(m);
  if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

// x[n] = (Dinv^(1/2) g / s, f)
__global__ void k_features(const float* __restrict__ g, VLay lay, long long nodes, const double* scale,
                           const float* __restrict__ dinv, const float* __restrict__ feat, float* __restrict__ x) {
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (node >= nodes) return;
  const double s = scale[n] > 0 ? scale[n] : 1.0;
  const float inv = (float)(1.0 / s);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[((long long)n * 6 + c) * nodes + node] = g[node * lay.ns + n * lay.rs + c] * sqrtf(dinv[node * 3 + c]) * inv;
    x[((long long)n * 6 + 3 + c) * nodes + node] = feat[c * nodes + node];
  }
}

// delta = Dinv g + Dinv^(1/2) s u
__global__ void k_delta(const float* __restrict__ g, VLay lay, long long nodes, const double* scale,
                        const float* __restrict__ u, const float* __restrict__ dinv, float* __restrict__ delta,
                        const int* active) {
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (node >= nodes || (active && !active[n])) return;
  const float s = (float)(scale[n] > 0 ? scale[n] : 1.0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long o = node * lay.ns + n * lay.rs + c;
    const float di = dinv[node * 3 + c];
    delta[o] = di * g[o] + sqrtf(di) * s * u[((long long)n * 3 + c) * nodes + node];
  }
}

// du = (A delta - g) Dinv^(1/2) s / norm; loss partials (0.5 delta.A delta - delta.g) / norm
__global__ __launch_bounds__(256) void k_energy_grad(const float* __restrict__ g, const float* __restrict__ delta,
                                                     const float* __restrict__ ad, VLay lay, long long nodes,
                                                     const double* scale, const double* norm,
                                                     const float* __restrict__ dinv, float* __restrict__ du,
                                                     double* partial) {
  const int n = blockIdx.y;
  const double s = scale[n] > 0 ? scale[n] : 1.0;
  const double nrm = norm[n] > 0 ? norm[n] : 1.0;
  double acc = 0.0;
  for (long long node = (long long)blockIdx.x * 256 + threadIdx.x; node < nodes; node += (long long)gridDim.x * 256) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long o = node * lay.ns + n * lay.rs + c;
      const double gv = g[o], dv = delta[o], av = ad[o];
      acc += (0.5 * dv * av - dv * gv) / nrm;
      du[((long long)n * 3 + c) * nodes + node] = (float)((av - gv) * sqrt((double)dinv[node * 3 + c]) * s / nrm);
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[(long long)n * gridDim.x + blockIdx.x] = acc;
}

__global__ void k_omega(const double* dg, const double* dad, double* omega, int n, const int* active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool on = !active || active[i];
  omega[i] = (on && dg[i] > 0 && dad[i] > 0) ? dg[i] / dad[i] : 0.0;
}

__global__ void k_scale_samples(float* __restrict__ v, VLay lay, long long nodes, const double* omega,
                                const int* active) {
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (node >= nodes || (active && !active[n])) return;
  const float w = (float)omega[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) v[node * lay.ns + n * lay.rs + c] *= w;
}

// dst[(base + n) * 3 * nodes + node * 3 + c] = src (layout lay), zero for inactive columns
__global__ void k_gather_samples(const float* __restrict__ src, VLay lay, long long nodes, const int* active,
                                 float* __restrict__ dst) {
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (node >= nodes) return;
  const bool on = !active || active[n];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    dst[((long long)n * nodes + node) * 3 + c] = on ? src[node * lay.ns + n * lay.rs + c] : 0.f;
}

// ---- network --------------------------------------------------------------------------------
struct NetLayer {
  const char* name;
  int kind;  // 0: conv stride 1, 1: conv stride 2, 2: up (adjoint of stride 2)
  int cin, cout;
  long long woff, boff;
};

static const NetLayer kLayers[] = {
    {"enc1a", 0, 6, 16, 0, 0},  {"enc1b", 0, 16, 16, 0, 0}, {"enc2a", 1, 16, 32, 0, 0}, {"enc2b", 0, 32, 32, 0, 0},
    {"enc3a", 1, 32, 64, 0, 0}, {"enc3b", 0, 64, 64, 0, 0}, {"enc4a", 1, 64, 64, 0, 0}, {"enc4b", 0, 64, 64, 0, 0},
    {"up3", 2, 64, 64, 0, 0},   {"dec3", 0, 64, 32, 0, 0},  {"up2", 2, 32, 32, 0, 0},   {"dec2", 0, 32, 16, 0, 0},
    {"up1", 2, 16, 16, 0, 0},   {"out", 0, 16, 3, 0, 0},
};
constexpr int kNumLayers = sizeof(kLayers) / sizeof(kLayers[0]);
enum {
  L_ENC1A, L_ENC1B, L_ENC2A, L_ENC2B, L_ENC3A, L_ENC3B, L_ENC4A, L_ENC4B, L_UP3, L_DEC3, L_UP2, L_DEC2, L_UP1, L_OUT
};
enum {
  A_X, A_E1A, A_S1, A_E2A, A_S2, A_E3A, A_S3, A_E4A, A_B4, A_U3, A_D3, A_U2, A_D2, A_U1, A_OUT, A_COUNT
};

struct NetBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t n) {
    if (n <= bytes) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    DFMI_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  template <typename T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
  ~NetBuf() {
    if (p) (void)hipFree(p);
  }
};

struct NetImpl {
  NetLayer layers[kNumLayers];
  long long nparam = 0;
  NetBuf master, bf, grad, m, v, wpartial;
  NetBuf act[A_COUNT];
  NetBuf gA, gB, gC, gS1, gS2, gS3;  // gradient work
  NetBuf feat, scal, part, dmaxb, du, delta, ad;
  NetDims dims[4]{};
  int nsamples = 0;
  int steps = 0;
  hipStream_t stream = nullptr;
  const void* feat_key = nullptr;
  long long feat_nodes = 0;

  NetImpl() {
    long long off = 0;
    for (int l = 0; l < kNumLayers; ++l) {
      layers[l] = kLayers[l];
      layers[l].woff = off;
      off += (long long)layers[l].cin * layers[l].cout * 27;
      layers[l].boff = off;
      off += layers[l].cout;
    }
    nparam = off;
    master.ensure(sizeof(float) * nparam);
    bf.ensure(sizeof(uint16_t) * nparam);
    grad.ensure(sizeof(float) * nparam);
    m.ensure(sizeof(float) * nparam);
    v.ensure(sizeof(float) * nparam);
    DFMI_HIP(hipMemset(master.p, 0, sizeof(float) * nparam));
    DFMI_HIP(hipMemset(bf.p, 0, sizeof(uint16_t) * nparam));
    DFMI_HIP(hipMemset(grad.p, 0, sizeof(float) * nparam));
    reset_adam();
  }
  void reset_adam() {
    DFMI_HIP(hipMemset(m.p, 0, sizeof(float) * nparam));
    DFMI_HIP(hipMemset(v.p, 0, sizeof(float) * nparam));
    steps = 0;
  }
  const uint16_t* W(int l) const { return bf.as<uint16_t>() + layers[l].woff; }
  const float* B(int l) const { return master.as<float>() + layers[l].boff; }
  float* GW(int l) { return grad.as<float>() + layers[l].woff; }
  float* GB(int l) { return grad.as<float>() + layers[l].boff; }
  // output dims of each activation
  NetDims dim_of(int a) const {
    switch (a) {
      case A_X: case A_E1A: case A_S1: case A_U1: case A_OUT: return dims[0];
      case A_E2A: case A_S2: case A_U2: case A_D2: return dims[1];
      case A_E3A: case A_S3: case A_U3: case A_D3: return dims[2];
      default: return dims[3];
    }
  }
  int chan_of(int a) const {
    switch (a) {
      case A_X: return 6;
      case A_E1A: case A_S1: case A_U1: case A_D2: return 16;
      case A_E2A: case A_S2: case A_U2: case A_D3: return 32;
      case A_OUT: return 3;
      default: return 64;
    }
  }
  void prepare(int n, const NetDims& d1) {
    nsamples = n;
    dims[0] = d1;
    for (int l = 1; l < 4; ++l)
      dims[l] = NetDims{(dims[l - 1].x - 1) / 2 + 1, (dims[l - 1].y - 1) / 2 + 1, (dims[l - 1].z - 1) / 2 + 1};
    for (int a = 0; a < A_COUNT; ++a) act[a].ensure(sizeof(float) * (size_t)n * chan_of(a) * dim_of(a).n());
  }
};

static void launch_conv(NetImpl* N, int stride, const float* in, NetDims din, int ci, const uint16_t* w, int co,
                        const float* bias, const float* add, const float* act, int epi, float* out, NetDims dout) {
  const dim3 grid((unsigned)((dout.n() + 255) / 256), (unsigned)(N->nsamples * ((co + CT - 1) / CT)));
  if (stride == 1)
    hipLaunchKernelGGL(k_conv<1>, grid, dim3(256), 0, N->stream, in, din, ci, w, co, bias, add, act, epi, out, dout);
  else
    hipLaunchKernelGGL(k_conv<2>, grid, dim3(256), 0, N->stream, in, din, ci, w, co, bias, add, act, epi, out, dout);
  DFMI_HIP(hipGetLastError());
}

static void launch_adj(NetImpl* N, int stride, const float* in, NetDims din, int no, const uint16_t* w, int ni,
                       const float* bias, const float* add, const float* act, int epi, float* out, NetDims dout) {
  const dim3 grid((unsigned)((dout.n() + 255) / 256), (unsigned)(N->nsamples * ((ni + CT - 1) / CT)));
  if (stride == 1)
    hipLaunchKernelGGL(k_conv_adj<1>, grid, dim3(256), 0, N->stream, in, din, no, w, ni, bias, add, act, epi, out,
                       dout);
  else
    hipLaunchKernelGGL(k_conv_adj<2>, grid, dim3(256), 0, N->stream, in, din, no, w, ni, bias, add, act, epi, out,
                       dout);
  DFMI_HIP(hipGetLastError());
}

// dW[o][i][t] for y = conv_S(x) with dy over the y grid; db[o] if db != nullptr
static void launch_wgrad(NetImpl* N, int stride, const float* x, NetDims dx, int ni, const float* dy, NetDims dyd,
                         int no, float* dw, float* db, const float* dbsrc, NetDims dbd, int dbc) {
  const long long total = (long long)N->nsamples * dyd.n();
  const int chunks = (int)std::max(1LL, std::min(64LL, (total + 2047) / 2048));
  const long long per = (total + chunks - 1) / chunks;
  const long long nw = (long long)no * ni * 27;
  N->wpartial.ensure(sizeof(float) * (size_t)chunks * nw);
  const dim3 grid(chunks, (no + WG_O - 1) / WG_O, (ni + WG_I - 1) / WG_I);
  if (stride == 1)
    hipLaunchKernelGGL(k_wgrad<1>, grid, dim3(WG_I * 27), 0, N->stream, x, dx, ni, dy, dyd, no, N->nsamples, per,
                       N->wpartial.as<float>());
  else
    hipLaunchKernelGGL(k_wgrad<2>, grid, dim3(WG_I * 27), 0, N->stream, x, dx, ni, dy, dyd, no, N->nsamples, per,
                       N->wpartial.as<float>());
  DFMI_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, N->stream,
                     N->wpartial.as<float>(), chunks, nw, dw);
  DFMI_HIP(hipGetLastError());
  if (db) {
    hipLaunchKernelGGL(k_bias_grad, dim3(dbc), dim3(256), 0, N->stream, dbsrc, dbd.n(), dbc, N->nsamples, db);
    DFMI_HIP(hipGetLastError());
  }
}

void net_forward(NetImpl* N, int n, const NetDims& d1, const float* x) {
  N->prepare(n, d1);
  const NetDims* D = N->dims;
  auto A = [&](int a) { return N->act[a].as<float>(); };
  if (x != A(A_X))
    DFMI_HIP(hipMemcpyAsync(A(A_X), x, sizeof(float) * (size_t)n * 6 * D[0].n(), hipMemcpyDeviceToDevice,
                            N->stream));
  launch_conv(N, 1, A(A_X), D[0], 6, N->W(L_ENC1A), 16, N->B(L_ENC1A), nullptr, nullptr, EPI_RELU, A(A_E1A), D[0]);
  launch_conv(N, 1, A(A_E1A), D[0], 16, N->W(L_ENC1B), 16, N->B(L_ENC1B), nullptr, nullptr, EPI_RELU, A(A_S1), D[0]);
  launch_conv(N, 2, A(A_S1), D[0], 16, N->W(L_ENC2A), 32, N->B(L_ENC2A), nullptr, nullptr, EPI_RELU, A(A_E2A), D[1]);
  launch_conv(N, 1, A(A_E2A), D[1], 32, N->W(L_ENC2B), 32, N->B(L_ENC2B), nullptr, nullptr, EPI_RELU, A(A_S2), D[1]);
  launch_conv(N, 2, A(A_S2), D[1], 32, N->W(L_ENC3A), 64, N->B(L_ENC3A), nullptr, nullptr, EPI_RELU, A(A_E3A), D[2]);
  launch_conv(N, 1, A(A_E3A), D[2], 64, N->W(L_ENC3B), 64, N->B(L_ENC3B), nullptr, nullptr, EPI_RELU, A(A_S3), D[2]);
  launch_conv(N, 2, A(A_S3), D[2], 64, N->W(L_ENC4A), 64, N->B(L_ENC4A), nullptr, nullptr, EPI_RELU, A(A_E4A), D[3]);
  launch_conv(N, 1, A(A_E4A), D[3], 64, N->W(L_ENC4B), 64, N->B(L_ENC4B), nullptr, nullptr, EPI_RELU, A(A_B4), D[3]);
  launch_adj(N, 2, A(A_B4), D[3], 64, N->W(L_UP3), 64, N->B(L_UP3), A(A_S3), nullptr, EPI_RELU, A(A_U3), D[2]);
  launch_conv(N, 1, A(A_U3), D[2], 64, N->W(L_DEC3), 32, N->B(L_DEC3), nullptr, nullptr, EPI_RELU, A(A_D3), D[2]);
  launch_adj(N, 2, A(A_D3), D[2], 32, N->W(L_UP2), 32, N->B(L_UP2), A(A_S2), nullptr, EPI_RELU, A(A_U2), D[1]);
  launch_conv(N, 1, A(A_U2), D[1], 32, N->W(L_DEC2), 16, N->B(L_DEC2), nullptr, nullptr, EPI_RELU, A(A_D2), D[1]);
  launch_adj(N, 2, A(A_D2), D[1], 16, N->W(L_UP1), 16, N->B(L_UP1), A(A_S1), nullptr, EPI_RELU, A(A_U1), D[0]);
  launch_conv(N, 1, A(A_U1), D[0], 16, N->W(L_OUT), 3, N->B(L_OUT), nullptr, nullptr, EPI_NONE, A(A_OUT), D[0]);
}

const float* net_output(NetImpl* N) { return N->act[A_OUT].as<float>(); }

void net_copy_output(NetImpl* N, float* dst, long long count) {
  DFMI_HIP(hipMemcpyAsync(dst, N->act[A_OUT].p, sizeof(float) * count, hipMemcpyDeviceToDevice, N->stream));
  DFMI_HIP(hipStreamSynchronize(N->stream));
}

void net_backward(NetImpl* N, const float* dout) {
  const NetDims* D = N->dims;
  const int n = N->nsamples;
  auto A = [&](int a) { return N->act[a].as<float>(); };
  const size_t top = sizeof(float) * (size_t)n * 64 * D[0].n();
  N->gA.ensure(top);
  N->gB.ensure(top);
  N->gS1.ensure(sizeof(float) * (size_t)n * 16 * D[0].n());
  N->gS2.ensure(sizeof(float) * (size_t)n * 32 * D[1].n());
  N->gS3.ensure(sizeof(float) * (size_t)n * 64 * D[2].n());
  float* ga = N->gA.as<float>();
  float* gb = N->gB.as<float>();
  float* gs1 = N->gS1.as<float>();
  float* gs2 = N->gS2.as<float>();
  float* gs3 = N->gS3.as<float>();
  // out: y = conv1(u1)
  launch_wgrad(N, 1, A(A_U1), D[0], 16, dout, D[0], 3, N->GW(L_OUT), N->GB(L_OUT), dout, D[0], 3);
  launch_adj(N, 1, dout, D[0], 3, N->W(L_OUT), 16, nullptr, nullptr, A(A_U1), EPI_MASK, gs1, D[0]);  // d pre(u1)
  // up1: u1 = relu(adj(d2) + b + s1)
  launch_wgrad(N, 2, gs1, D[0], 16, A(A_D2), D[1], 16, N->GW(L_UP1), N->GB(L_UP1), gs1, D[0], 16);
  launch_conv(N, 2, gs1, D[0], 16, N->W(L_UP1), 16, nullptr, nullptr, A(A_D2), EPI_MASK, ga, D[1]);  // d pre(d2)
  // dec2: d2 = relu(conv1(u2))
  launch_wgrad(N, 1, A(A_U2), D[1], 32, ga, D[1], 16, N->GW(L_DEC2), N->GB(L_DEC2), ga, D[1], 16);
  launch_adj(N, 1, ga, D[1], 16, N->W(L_DEC2), 32, nullptr, nullptr, A(A_U2), EPI_MASK, gs2, D[1]);  // d pre(u2)
  // up2
  launch_wgrad(N, 2, gs2, D[1], 32, A(A_D3), D[2], 32, N->GW(L_UP2), N->GB(L_UP2), gs2, D[1], 32);
  launch_conv(N, 2, gs2, D[1], 32, N->W(L_UP2), 32, nullptr, nullptr, A(A_D3), EPI_MASK, ga, D[2]);  // d pre(d3)
  // dec3
  launch_wgrad(N, 1, A(A_U3), D[2], 64, ga, D[2], 32, N->GW(L_DEC3), N->GB(L_DEC3), ga, D[2], 32);
  launch_adj(N, 1, ga, D[2], 32, N->W(L_DEC3), 64, nullptr, nullptr, A(A_U3), EPI_MASK, gs3, D[2]);  // d pre(u3)
  // up3
  launch_wgrad(N, 2, gs3, D[2], 64, A(A_B4), D[3], 64, N->GW(L_UP3), N->GB(L_UP3), gs3, D[2], 64);
  launch_conv(N, 2, gs3, D[2], 64, N->W(L_UP3), 64, nullptr, nullptr, A(A_B4), EPI_MASK, ga, D[3]);  // d pre(b4)
  // enc4b: b4 = relu(conv1(e4a))
  launch_wgrad(N, 1, A(A_E4A), D[3], 64, ga, D[3], 64, N->GW(L_ENC4B), N->GB(L_ENC4B), ga, D[3], 64);
  launch_adj(N, 1, ga, D[3], 64, N->W(L_ENC4B), 64, nullptr, nullptr, A(A_E4A), EPI_MASK, gb, D[3]);
  // enc4a: e4a = relu(conv2(s3)); s3 also feeds u3
  launch_wgrad(N, 2, A(A_S3), D[2], 64, gb, D[3], 64, N->GW(L_ENC4A), N->GB(L_ENC4A), gb, D[3], 64);
  launch_adj(N, 2, gb, D[3], 64, N->W(L_ENC4A), 64, nullptr, gs3, A(A_S3), EPI_MASK, ga, D[2]);  // d pre(s3)
  // enc3b
  launch_wgrad(N, 1, A(A_E3A), D[2], 64, ga, D[2], 64, N->GW(L_ENC3B), N->GB(L_ENC3B), ga, D[2], 64);
  launch_adj(N, 1, ga, D[2], 64, N->W(L_ENC3B), 64, nullptr, nullptr, A(A_E3A), EPI_MASK, gb, D[2]);
  // enc3a
  launch_wgrad(N, 2, A(A_S2), D[1], 32, gb, D[2], 64, N->GW(L_ENC3A), N->GB(L_ENC3A), gb, D[2], 64);
  launch_adj(N, 2, gb, D[2], 64, N->W(L_ENC3A), 32, nullptr, gs2, A(A_S2), EPI_MASK, ga, D[1]);  // d pre(s2)
  // enc2b
  launch_wgrad(N, 1, A(A_E2A), D[1], 32, ga, D[1], 32, N->GW(L_ENC2B), N->GB(L_ENC2B), ga, D[1], 32);
  launch_adj(N, 1, ga, D[1], 32, N->W(L_ENC2B), 32, nullptr, nullptr, A(A_E2A), EPI_MASK, gb, D[1]);
  // enc2a
  launch_wgrad(N, 2, A(A_S1), D[0], 16, gb, D[1], 32, N->GW(L_ENC2A), N->GB(L_ENC2A), gb, D[1], 32);
  launch_adj(N, 2, gb, D[1], 32, N->W(L_ENC2A), 16, nullptr, gs1, A(A_S1), EPI_MASK, ga, D[0]);  // d pre(s1)
  // enc1b
  launch_wgrad(N, 1, A(A_E1A), D[0], 16, ga, D[0], 16, N->GW(L_ENC1B), N->GB(L_ENC1B), ga, D[0], 16);
  launch_adj(N, 1, ga, D[0], 16, N->W(L_ENC1B), 16, nullptr, nullptr, A(A_E1A), EPI_MASK, gb, D[0]);
  // enc1a
  launch_wgrad(N, 1, A(A_X), D[0], 6, gb, D[0], 16, N->GW(L_ENC1A), N->GB(L_ENC1A), gb, D[0], 16);
}

void net_adam(NetImpl* N, double lr, double b1, double b2, double eps) {
  N->steps += 1;
  const double c1 = 1.0 / (1.0 - std::pow(b1, N->steps)), c2 = 1.0 / (1.0 - std::pow(b2, N->steps));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((N->nparam + 255) / 256)), dim3(256), 0, N->stream, N->master.as<float>(),
                     N->grad.as<float>(), N->m.as<float>(), N->v.as<float>(), N->bf.as<uint16_t>(), N->nparam,
                     (float)lr, (float)b1, (float)b2, (float)eps, (float)c1, (float)c2);
  DFMI_HIP(hipGetLastError());
}

NetImpl* net_new() { return new NetImpl(); }
void net_delete(NetImpl* N) { delete N; }
void net_set_stream(NetImpl* N, hipStream_t s) { N->stream = s; }
long long net_param_count(const NetImpl* N) { return N->nparam; }

void net_set_params(NetImpl* N, const float* src) {
  DFMI_HIP(hipMemcpyAsync(N->master.p, src, sizeof(float) * N->nparam, hipMemcpyDeviceToDevice, N->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((N->nparam + 255) / 256)), dim3(256), 0, N->stream,
                     N->master.as<float>(), N->bf.as<uint16_t>(), N->nparam);
  DFMI_HIP(hipGetLastError());
  DFMI_HIP(hipStreamSynchronize(N->stream));
  N->reset_adam();
}

void net_get(NetImpl* N, int which, float* dst) {
  const NetBuf& b = which == 0 ? N->master : N->grad;
  DFMI_HIP(hipMemcpyAsync(dst, b.p, sizeof(float) * N->nparam, hipMemcpyDeviceToDevice, N->stream));
}

// ---- the correction on a level-1 multi-vector ---------------------------------------------------
// scalar workspace: 8 slices of w = max(n, 64) doubles: scale, -, delta.g, delta.A delta, omega, norm, loss, -
static int scal_width(int n) { return std::max(n, 64); }
static void ensure_scal(NetImpl* N, int n) {
  N->scal.ensure(sizeof(double) * (size_t)scal_width(n) * 8);
  N->part.ensure(sizeof(double) * (size_t)scal_width(n) * 64);
}

static void sample_scale(NetImpl* N, const float* g, VLay lay, long long nodes, int n, const float* dinv) {
  const int nb = (int)std::min<long long>(64, (nodes * 3 + 255) / 256);
  ensure_scal(N, n);
  hipLaunchKernelGGL(k_sample_absmax, dim3(nb, n), dim3(256), 0, N->stream, g, lay, nodes, dinv,
                     N->part.as<double>());
  DFMI_HIP(hipGetLastError());
  reduce_max(N->part.as<double>(), nb, n, N->scal.as<double>(), N->stream);
}

static void sample_dot(NetImpl* N, const float* a, const float* b, const float* w, VLay lay, long long nodes, int n,
                       double* out) {
  const int nb = (int)std::min<long long>(64, (nodes * 3 + 255) / 256);
  ensure_scal(N, n);
  hipLaunchKernelGGL(k_sample_dot, dim3(nb, n), dim3(256), 0, N->stream, a, b, w, lay, nodes, N->part.as<double>());
  DFMI_HIP(hipGetLastError());
  reduce_partials(N->part.as<double>(), nb, n, out, N->stream);
}

static void stiffness_feature(NetImpl* N, const float* dinv, long long nodes) {
  if (N->feat_key == dinv && N->feat_nodes == nodes) return;
  N->feat.ensure(sizeof(float) * 3 * nodes);
  ensure_scal(N, 1);
  N->dmaxb.ensure(sizeof(double));
  const int nb = (int)std::min<long long>(64, (nodes * 3 + 255) / 256);
  hipLaunchKernelGGL(k_min_positive_partial, dim3(nb), dim3(256), 0, N->stream, dinv, nodes * 3,
                     N->part.as<double>());
  DFMI_HIP(hipGetLastError());
  double* dmax = N->dmaxb.as<double>();
  reduce_max(N->part.as<double>(), nb, 1, dmax, N->stream);
  hipLaunchKernelGGL(k_stiffness_feature, dim3((unsigned)((nodes * 3 + 255) / 256)), dim3(256), 0, N->stream, dinv,
                     nodes, dmax, N->feat.as<float>());
  DFMI_HIP(hipGetLastError());
  N->feat_key = dinv;
  N->feat_nodes = nodes;
}

void net_invalidate_feature(NetImpl* N) { N->feat_key = nullptr; }

// delta = Dinv g + Dinv^(1/2) s N(Dinv^(1/2) g / s, f) for n samples of level-1 geometry g1 in layout lay
void net_correct(NetImpl* N, const Geom& g1, const float* g, VLay lay, int n, const float* dinv, const int* active,
                 float* delta) {
  const long long nodes = g1.nodes();
  stiffness_feature(N, dinv, nodes);
  sample_scale(N, g, lay, nodes, n, dinv);
  const NetDims d1{g1.nx, g1.ny, g1.nz};
  N->prepare(n, d1);
  hipLaunchKernelGGL(k_features, dim3((unsigned)((nodes + 255) / 256), n), dim3(256), 0, N->stream, g, lay, nodes,
                     N->scal.as<double>(), dinv, N->feat.as<float>(), N->act[A_X].as<float>());
  DFMI_HIP(hipGetLastError());
  net_forward(N, n, d1, N->act[A_X].as<float>());
  hipLaunchKernelGGL(k_delta, dim3((unsigned)((nodes + 255) / 256), n), dim3(256), 0, N->stream, g, lay, nodes,
                     N->scal.as<double>(), N->act[A_OUT].as<float>(), dinv, delta, active);
  DFMI_HIP(hipGetLastError());
}

// omega[n] = (delta.g) / (delta.A delta) (0 when not a descent step); delta *= omega
void net_energy_step(NetImpl* N, const float* g, const float* ad, float* delta, VLay lay, long long nodes, int n,
                     const int* active) {
  ensure_scal(N, n);
  const int w = scal_width(n);
  double* dg = N->scal.as<double>() + 2 * (size_t)w;
  double* dad = dg + w;
  double* om = dad + w;
  sample_dot(N, delta, g, nullptr, lay, nodes, n, dg);
  sample_dot(N, delta, ad, nullptr, lay, nodes, n, dad);
  hipLaunchKernelGGL(k_omega, dim3((n + 63) / 64), dim3(64), 0, N->stream, dg, dad, om, n, active);
  DFMI_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scale_samples, dim3((unsigned)((nodes + 255) / 256), n), dim3(256), 0, N->stream, delta, lay,
                     nodes, om, active);
  DFMI_HIP(hipGetLastError());
}

// one training step on n samples g (column-major, level-1 geometry) with operator apply `ad = A delta`
double net_train_step(NetImpl* N, const Geom& g1, const float* g, int n, const float* dinv, double lr,
                      const std::function<void(const float*, float*)>& apply) {
  const long long nodes = g1.nodes();
  const VLay lay = column_major(g1);
  N->delta.ensure(sizeof(float) * (size_t)n * 3 * nodes);
  N->ad.ensure(sizeof(float) * (size_t)n * 3 * nodes);
  N->du.ensure(sizeof(float) * (size_t)n * 3 * nodes);
  net_correct(N, g1, g, lay, n, dinv, nullptr, N->delta.as<float>());
  apply(N->delta.as<float>(), N->ad.as<float>());
  const int w = scal_width(n);
  double* norm = N->scal.as<double>() + 5 * (size_t)w;
  double* loss = norm + w;
  sample_dot(N, g, g, dinv, lay, nodes, n, norm);
  const int nb = (int)std::min<long long>(64, (nodes + 255) / 256);
  hipLaunchKernelGGL(k_energy_grad, dim3(nb, n), dim3(256), 0, N->stream, g, N->delta.as<float>(),
                     N->ad.as<float>(), lay, nodes, N->scal.as<double>(), norm, dinv, N->du.as<float>(),
                     N->part.as<double>());
  DFMI_HIP(hipGetLastError());
  reduce_partials(N->part.as<double>(), nb, n, loss, N->stream);
  net_backward(N, N->du.as<float>());
  if (lr > 0) net_adam(N, lr, 0.9, 0.999, 1e-8);
  std::vector<double> h(n);
  DFMI_HIP(hipMemcpyAsync(h.data(), loss, sizeof(double) * n, hipMemcpyDeviceToHost, N->stream));
  DFMI_HIP(hipStreamSynchronize(N->stream));
  double s = 0;
  for (double x : h) s += x;
  return s;
}

void gather_samples(const float* src, VLay lay, long long nodes, int n, const int* active, float* dst,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_gather_samples, dim3((unsigned)((nodes + 255) / 256), n), dim3(256), 0, s, src, lay, nodes,
                     active, dst);
  DFMI_HIP(hipGetLastError());
}

}  // namespace dfmi
