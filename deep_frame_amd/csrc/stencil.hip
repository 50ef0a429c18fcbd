// Coarse levels: Galerkin operators stored as assembled 27-point 3x3-block stencils
// (structure of arrays st[(d*9 + 3c + c') * nodes + n], d = 9(dx+1) + 3(dy+1) + (dz+1)),
// the stencil form of the per-coarse-element Galerkin blocks of
// deep_frame/topology_multigrid.py:174-181 (A_c = P^T M A M P), plus the transfer
// operators (:314-316), the coarsest dense solve (:230-283 uses a sparse direct
// factorisation; here a dense inverse applied as one product per V-cycle).
#include "dfmi_internal.h"

namespace dfmi {

__host__ __device__ inline bool live_bit(uint8_t m, int c) { return (m >> c) & 1; }
__device__ inline double weight1(int t) { return t == 0 ? 1.0 : ((t == 1 || t == -1) ? 0.5 : 0.0); }

// Thread per (node, pair of right-hand sides), right-hand sides fastest: with the interleaved layout the
// threads of one node read one contiguous run per neighbour.  Thread g of a node takes columns g and
// g + groups, so each load instruction of the node's threads covers one dense run of columns (no
// interleaved gaps that would make the address unit visit every cache line twice).  The node's 27 x 9 stencil entries are
// staged once per workgroup in LDS and broadcast to its threads.  No bounds or mask tests in the
// gather: stencil entries towards missing or dead neighbours are zero (Galerkin product and
// stencil_mask) and level vectors are zero on dead DOFs, so a neighbour index that runs off a grid
// line only multiplies a zero; it is clamped into the array.
constexpr int SNT = 512;
constexpr int SP = 244;  // padded LDS row of 243 stencil entries


template <typename T, int RP, int MODE>
__global__ __launch_bounds__(SNT) void k_stencil_apply(Geom g, const T* __restrict__ st,
                                                        const uint8_t* __restrict__ mask, ApplyArgs a, VLay lay,
                                                        int nodes_per_block, int groups) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sst = reinterpret_cast<T*>(smem_raw);  // [nodes_per_block][SP]
  const int N = (int)g.nodes();
  const int n0 = (int)xcd_block(blockIdx.x, gridDim.x) * nodes_per_block;
  const int t = threadIdx.x;
  {  // stage the block's stencils: thread t handles node j = t % npb and entries e = t / npb, + SNT / npb, ...
    const int npb = nodes_per_block, step = SNT / npb;
    const int j = t % npb, e0 = t / npb;
    const int n = n0 + j;
    if (e0 < step) {
      const T* src = st + n;
      for (int e = e0; e < 243; e += step) sst[j * SP + e] = n < N ? src[(long long)e * N] : T(0);
    }
  }
  __syncthreads();
  const int j = t / groups, grp = t - (t / groups) * groups;
  const int n = n0 + j;
  if (j >= nodes_per_block || n >= N) return;
  int col[RP];
  bool on[RP];
#pragma unroll
  for (int q = 0; q < RP; ++q) {
    col[q] = grp + q * groups;
    on[q] = col[q] < a.k && (!a.active || a.active[col[q]]);
  }
  bool any = false;
#pragma unroll
  for (int q = 0; q < RP; ++q) any = any || on[q];
  if (!any) return;
  const int ns = (int)lay.ns, rs = (int)lay.rs;
  const T* X = reinterpret_cast<const T*>(a.x);
  // the buffer spans the whole multi-vector (interleaved or column-major)
  const long long extent = ((long long)(N - 1) * ns + (long long)(a.k - 1) * rs + 3) * (long long)sizeof(T);
  const __amdgpu_buffer_rsrc_t xr =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(X), (short)0, (int)extent, kBufferWord3);
  const T* s = sst + j * SP;
  T acc[RP][3];
#pragma unroll
  for (int q = 0; q < RP; ++q) acc[q][0] = acc[q][1] = acc[q][2] = T(0);
  const int py = g.nz, px = g.ny * g.nz;
  const int nsb = ns * (int)sizeof(T);
  unsigned cb[RP];  // byte offset of each of this thread's columns within node n
#pragma unroll
  for (int q = 0; q < RP; ++q) cb[q] = (unsigned)(n * nsb + min(col[q], a.k - 1) * rs * (int)sizeof(T));
#pragma unroll 3
  for (int d = 0; d < 27; ++d) {
    const int od = ((d / 9 - 1) * px + ((d / 3) % 3 - 1) * py + (d % 3 - 1)) * nsb;  // uniform
    const T* sd = s + d * 9;
    const T s0 = sd[0], s1 = sd[1], s2 = sd[2], s3 = sd[3], s4 = sd[4], s5 = sd[5], s6 = sd[6], s7 = sd[7], s8 = sd[8];
#pragma unroll
    for (int q = 0; q < RP; ++q) {
      T x0, x1, x2;  // a missing odd column re-reads the last one
      load3(xr, cb[q] + (unsigned)od, x0, x1, x2);
      acc[q][0] = fma(s0, x0, fma(s1, x1, fma(s2, x2, acc[q][0])));
      acc[q][1] = fma(s3, x0, fma(s4, x1, fma(s5, x2, acc[q][1])));
      acc[q][2] = fma(s6, x0, fma(s7, x1, fma(s8, x2, acc[q][2])));
    }
  }
  // epilogue: the node's three components of a column are one load or store per operand (the result,
  // b and x_prev share the operand's layout and extent)
  const uint8_t mn = mask[n];
  const __amdgpu_buffer_rsrc_t yr = buffer(a.y, extent);
  T dv[3] = {T(0), T(0), T(0)};
  if (MODE == MODE_CHEB) load3(buffer(a.dinv, (long long)N * 3 * (long long)sizeof(T)), (unsigned)(n * 3 * (int)sizeof(T)),
                               dv[0], dv[1], dv[2]);
#pragma unroll
  for (int q = 0; q < RP; ++q) {
    if (!on[q]) continue;
    const unsigned ob = (unsigned)((n * ns + col[q] * rs) * (int)sizeof(T));
    T y[3];
    if (MODE == MODE_APPLY) {
      y[0] = acc[q][0];
      y[1] = acc[q][1];
      y[2] = acc[q][2];
    } else if (MODE == MODE_RESID) {
      T b[3];
      load3(buffer(a.b, extent), ob, b[0], b[1], b[2]);
#pragma unroll
      for (int c = 0; c < 3; ++c) y[c] = b[c] - acc[q][c];
    } else if (MODE == MODE_CHEB) {
      T x[3], xp[3] = {T(0), T(0), T(0)}, b[3];
      load3(xr, ob, x[0], x[1], x[2]);
      if (a.xp) load3(buffer(a.xp, extent), ob, xp[0], xp[1], xp[2]);
      load3(buffer(a.b, extent), ob, b[0], b[1], b[2]);
#pragma unroll
      for (int c = 0; c < 3; ++c) y[c] = x[c] + (T)a.c1 * (x[c] - xp[c]) + (T)a.c2 * dv[c] * (b[c] - acc[q][c]);
    }
    store3(yr, ob, live_bit(mn, 0) ? y[0] : T(0), live_bit(mn, 1) ? y[1] : T(0), live_bit(mn, 2) ? y[2] : T(0));
  }
}

// One column (the power iterations of the set-up): a thread per node reads its 243 stencil entries straight
// from the structure-of-arrays stencil (coalesced across the wavefront) and gathers its 27 neighbours; the
// workgroup-staged form above would leave all but one thread per node idle.  Neighbour offsets that leave
// the array read zeros through the buffer's range check; those that wrap around a grid line meet zero entries.
template <typename T>
__global__ __launch_bounds__(256) void k_stencil_one(Geom g, const T* __restrict__ st, const uint8_t* __restrict__ mask,
                                                     const T* __restrict__ x, T* __restrict__ y) {
  const int N = (int)g.nodes();
  const int n = (int)xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (n >= N) return;
  const __amdgpu_buffer_rsrc_t xr = buffer(x, (long long)N * 3 * (long long)sizeof(T));
  const int py = g.nz, px = g.ny * g.nz;
  T acc[3] = {T(0), T(0), T(0)};
#pragma unroll 3
  for (int d = 0; d < 27; ++d) {
    const int od = (d / 9 - 1) * px + ((d / 3) % 3 - 1) * py + (d % 3 - 1);
    T x0, x1, x2;
    load3(xr, (unsigned)((n + od) * 3 * (int)sizeof(T)), x0, x1, x2);
    const T* sd = st + (long long)d * 9 * N + n;
    acc[0] = fma(sd[0], x0, fma(sd[(long long)N], x1, fma(sd[2LL * N], x2, acc[0])));
    acc[1] = fma(sd[3LL * N], x0, fma(sd[4LL * N], x1, fma(sd[5LL * N], x2, acc[1])));
    acc[2] = fma(sd[6LL * N], x0, fma(sd[7LL * N], x1, fma(sd[8LL * N], x2, acc[2])));
  }
  const uint8_t mn = mask[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) y[3LL * n + c] = live_bit(mn, c) ? acc[c] : T(0);
}

// right-hand sides per thread and nodes per workgroup (staged stencils kept within 48 KB of LDS)
// (2 per thread measured fastest at 42 right-hand sides on the level-1 grid of 256 x 256 x 48: 3 and 6
// per thread trade L2 re-reads for register pressure and lose)
static int stencil_rp(int k) { return k >= 4 ? 2 : 1; }

template <typename T>
static int stencil_npb(int k) {
  const int groups = (k + stencil_rp(k) - 1) / stencil_rp(k);
  const int cap = (48 * 1024) / (int)(SP * sizeof(T));
  return std::max(1, std::min(SNT / std::max(1, groups), cap));
}

int stencil_blocks(const Geom& g, int k) {
  const int npb = stencil_npb<float>(k);
  return (int)((g.nodes() + npb - 1) / npb);
}

template <typename T>
void stencil_apply(const Geom& g, const T* st, const uint8_t* mask, int mode, const ApplyArgs& a, VLay lay,
                   hipStream_t s, int* nblocks_out) {
  const int rp = stencil_rp(a.k), groups = (a.k + rp - 1) / rp;
  if (groups > SNT) throw std::runtime_error("stencil_apply: too many right-hand sides per call");
  if (g.nodes() * 3 * (long long)a.k * (long long)sizeof(T) > (long long)INT32_MAX)
    throw std::runtime_error("stencil_apply: level vectors too large for 32-bit byte offsets");
  if (a.k == 1 && mode == MODE_APPLY && !a.active && lay.ns == 3) {
    const unsigned b1 = (unsigned)((g.nodes() + 255) / 256);
    if (nblocks_out) *nblocks_out = (int)b1;
    hipLaunchKernelGGL(k_stencil_one<T>, dim3(b1), dim3(256), 0, s, g, st, mask, reinterpret_cast<const T*>(a.x),
                       reinterpret_cast<T*>(a.y));
    DFMI_HIP(hipGetLastError());
    return;
  }
  const int npb = stencil_npb<T>(a.k);
  const unsigned bx = (unsigned)((g.nodes() + npb - 1) / npb);
  const size_t lds = sizeof(T) * SP * (size_t)npb;
  if (nblocks_out) *nblocks_out = bx;
#define DFMI_SLAUNCH(RP, M) \
  hipLaunchKernelGGL((k_stencil_apply<T, RP, M>), dim3(bx), dim3(SNT), lds, s, g, st, mask, a, lay, npb, groups)
#define DFMI_SMODE(RP)                                                \
  switch (mode) {                                                     \
    case MODE_APPLY: DFMI_SLAUNCH(RP, MODE_APPLY); break;              \
    case MODE_RESID: DFMI_SLAUNCH(RP, MODE_RESID); break;              \
    case MODE_CHEB: DFMI_SLAUNCH(RP, MODE_CHEB); break;                \
    default: throw std::runtime_error("stencil_apply: bad mode");     \
  }
  if (rp == 2) {
    DFMI_SMODE(2)
  } else {
    DFMI_SMODE(1)
  }
#undef DFMI_SMODE
#undef DFMI_SLAUNCH
  DFMI_HIP(hipGetLastError());
}

template void stencil_apply<float>(const Geom&, const float*, const uint8_t*, int, const ApplyArgs&, VLay,
                                   hipStream_t, int*);
template void stencil_apply<double>(const Geom&, const double*, const uint8_t*, int, const ApplyArgs&, VLay,
                                    hipStream_t, int*);

// ---- Galerkin product P^T A P in stencil form -------------------------------------------
// One axis of the separable Galerkin product: P = P_x (x) P_y (x) P_z, so P^T A P is the 1-D Galerkin
// coarsening applied along x, then y, then z, each on both sides; the intermediate operators stay 27-point
// 3x3-block stencils.  Output node N, offset D: along the coarsened axis the sum over input nodes
// f = 2 N + e (weight w(e)) and input offsets df with f + df inside the support of N + D (weight
// w(f + df - 2 (N + D))); the other axes pass through (same node coordinate, same offset).  One thread per
// (output node, offset), nine entries each.
__global__ void k_rap_axis(Geom gi, const double* __restrict__ sin, Geom go, double* __restrict__ sout, int axis) {
  const long long No = go.nodes(), Ni = gi.nodes();
  const unsigned logical = xcd_block(blockIdx.x, gridDim.x);
  const int D = (int)(logical % 27);
  const long long n = (long long)(logical / 27) * blockDim.x + threadIdx.x;
  if (n >= No) return;
  int c[3] = {(int)(n / ((long long)go.nz * go.ny)), (int)((n / go.nz) % go.ny), (int)(n % go.nz)};
  const int dd[3] = {D / 9 - 1, (D / 3) % 3 - 1, D % 3 - 1};
  const int ni[3] = {gi.nx, gi.ny, gi.nz};
  const int mo = c[axis] + dd[axis];  // the coarse neighbour along the axis
  const int no_axis = axis == 0 ? go.nx : (axis == 1 ? go.ny : go.nz);
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (mo >= 0 && mo < no_axis) {
    for (int e = -1; e <= 1; ++e) {
      const int f = 2 * c[axis] + e;
      if (f < 0 || f >= ni[axis]) continue;
      const double we = e == 0 ? 1.0 : 0.5;
      for (int df = -1; df <= 1; ++df) {
        const int t = f + df - 2 * mo;
        if (t < -1 || t > 1) continue;
        if (f + df < 0 || f + df >= ni[axis]) continue;
        const double w = we * (t == 0 ? 1.0 : 0.5);
        int fi[3] = {c[0], c[1], c[2]}, od[3] = {dd[0], dd[1], dd[2]};
        fi[axis] = f;
        od[axis] = df;
        const long long fn = ((long long)fi[0] * gi.ny + fi[1]) * gi.nz + fi[2];
        const int din = 9 * (od[0] + 1) + 3 * (od[1] + 1) + (od[2] + 1);
        const double* src = sin + (long long)din * 9 * Ni + fn;
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[q] = fma(w, src[q * Ni], acc[q]);
      }
    }
  }
  for (int q = 0; q < 9; ++q) sout[((long long)D * 9 + q) * No + n] = acc[q];
}

void stencil_rap(const Geom& fine, const double* fst, const Geom& coarse, double* cst, hipStream_t s) {
  // fine (nx, ny, nz) -> (Nx, ny, nz) -> (Nx, Ny, nz) -> (Nx, Ny, Nz) = coarse
  Geom g1 = fine, g2 = fine;
  g1.nx = coarse.nx;
  g1.ex = coarse.ex;
  g2.nx = coarse.nx;
  g2.ex = coarse.ex;
  g2.ny = coarse.ny;
  g2.ey = coarse.ey;
  double *s1 = nullptr, *s2 = nullptr;
  DFMI_HIP(hipMallocAsync((void**)&s1, sizeof(double) * 243 * (size_t)g1.nodes(), s));
  DFMI_HIP(hipMallocAsync((void**)&s2, sizeof(double) * 243 * (size_t)g2.nodes(), s));
  const Geom gin[3] = {fine, g1, g2}, gout[3] = {g1, g2, coarse};
  const double* in[3] = {fst, s1, s2};
  double* out[3] = {s1, s2, cst};
  for (int a = 0; a < 3; ++a) {
    hipLaunchKernelGGL(k_rap_axis, dim3((unsigned)((gout[a].nodes() + 255) / 256) * 27), dim3(256), 0, s, gin[a],
                       in[a], gout[a], out[a], a);
    DFMI_HIP(hipGetLastError());
  }
  DFMI_HIP(hipFreeAsync(s1, s));
  DFMI_HIP(hipFreeAsync(s2, s));
}

__global__ void k_stencil_diag(Geom g, const double* __restrict__ st, double* __restrict__ diag) {
  const long long N = g.nodes();
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  for (int c = 0; c < 3; ++c) diag[n * 3 + c] = st[((long long)13 * 9 + 4 * c) * N + n];
}

void stencil_diagonal(const Geom& g, const double* st, double* diag, hipStream_t s) {
  hipLaunchKernelGGL(k_stencil_diag, dim3((unsigned)((g.nodes() + 255) / 256)), dim3(256), 0, s, g, st, diag);
  DFMI_HIP(hipGetLastError());
}

__global__ void k_dead_mask(Geom g, const double* __restrict__ diag, double thr, uint8_t* __restrict__ mask) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nodes()) return;
  uint8_t m = 0;
  for (int c = 0; c < 3; ++c)
    if (diag[n * 3 + c] > thr) m |= (uint8_t)(1 << c);
  mask[n] = m;
}

void dead_mask(const Geom& g, const double* diag, double threshold, uint8_t* mask, hipStream_t s) {
  hipLaunchKernelGGL(k_dead_mask, dim3((unsigned)((g.nodes() + 255) / 256)), dim3(256), 0, s, g, diag, threshold,
                     mask);
  DFMI_HIP(hipGetLastError());
}

// zero rows and columns of masked DOFs: st_nm[c][c'] *= M_nc M_mc'
__global__ void k_stencil_mask(Geom g, double* __restrict__ st, const uint8_t* __restrict__ mask) {
  const long long N = g.nodes();
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (n >= N) return;
  const int l = (int)(n % g.nz), j = (int)((n / g.nz) % g.ny), i = (int)(n / ((long long)g.nz * g.ny));
  const int mi = i + d / 9 - 1, mj = j + (d / 3) % 3 - 1, ml = l + d % 3 - 1;
  const bool in = mi >= 0 && mj >= 0 && ml >= 0 && mi < g.nx && mj < g.ny && ml < g.nz;
  const uint8_t mn = mask[n];
  const uint8_t mm = in ? mask[((long long)mi * g.ny + mj) * g.nz + ml] : 0;
  for (int c = 0; c < 3; ++c)
    for (int cc = 0; cc < 3; ++cc)
      if (!live_bit(mn, c) || !live_bit(mm, cc)) st[((long long)d * 9 + 3 * c + cc) * N + n] = 0.0;
}

void stencil_mask(const Geom& g, double* st, const uint8_t* mask, hipStream_t s) {
  hipLaunchKernelGGL(k_stencil_mask, dim3((unsigned)((g.nodes() + 255) / 256), 27), dim3(256), 0, s, g, st, mask);
  DFMI_HIP(hipGetLastError());
}

// ---- transfer operators ------------------------------------------------------------------
// thread id -> (node, column): column fastest for interleaved outputs, node fastest otherwise
// Transfers.  Column-major vectors: grid (node blocks, columns), so no division recovers the column;
// interleaved vectors: one thread per (node, column), column fastest.  Indices are 32-bit (levels are
// checked to hold fewer than 2^31 values) and sums are in the working precision (the weights are powers
// of two, exact in both precisions).
__device__ inline void node_xyz(int n, const Geom& g, int& i, int& j, int& l) {
  const int plane = g.ny * g.nz;
  i = n / plane;
  const int rem = n - i * plane;
  j = rem / g.nz;
  l = rem - j * g.nz;
}

template <typename T>
__device__ inline void restrict_node(const Geom& gf, const T* __restrict__ F, int fns, const Geom& gc, int n,
                                     T (&acc)[3]) {
  int I, J, K;
  node_xyz(n, gc, I, J, K);
  acc[0] = acc[1] = acc[2] = T(0);
#pragma unroll
  for (int e = 0; e < 27; ++e) {
    const int ex = e / 9 - 1, ey = (e / 3) % 3 - 1, ez = e % 3 - 1;
    const int fi = 2 * I + ex, fj = 2 * J + ey, fk = 2 * K + ez;
    if (fi < 0 || fj < 0 || fk < 0 || fi >= gf.nx || fj >= gf.ny || fk >= gf.nz) continue;
    const T w = T((ex == 0 ? 1.0 : 0.5) * (ey == 0 ? 1.0 : 0.5) * (ez == 0 ? 1.0 : 0.5));
    const T* v = F + ((fi * gf.ny + fj) * gf.nz + fk) * fns;
    acc[0] += w * v[0];
    acc[1] += w * v[1];
    acc[2] += w * v[2];
  }
}

template <typename T>
__device__ inline void prolong_node(const Geom& gc, const T* __restrict__ C, int cns, const Geom& gf, int n,
                                    T (&acc)[3]) {
  int i, j, kz;
  node_xyz(n, gf, i, j, kz);
  acc[0] = acc[1] = acc[2] = T(0);
  const int i0 = i >> 1, j0 = j >> 1, k0 = kz >> 1;
  // branch-free: always the eight parents of the coarse cell; an even coordinate takes its own parent with
  // weight 1 twice at half weight (the second read repeats the first), an odd one both neighbours at 1/2
  const int i1 = i0 + (i & 1), j1 = j0 + (j & 1), k1 = k0 + (kz & 1);
  const int pi[2] = {i0, i1}, pj[2] = {j0, j1}, pk[2] = {k0, k1};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const T* v = C + (((long long)pi[a] * gc.ny + pj[b]) * gc.nz + pk[d]) * cns;
        acc[0] += v[0];
        acc[1] += v[1];
        acc[2] += v[2];
      }
  acc[0] *= T(0.125);
  acc[1] *= T(0.125);
  acc[2] *= T(0.125);
}

template <typename T>
__global__ void k_restrict(Geom gf, const T* __restrict__ f, VLay fl, Geom gc, const uint8_t* __restrict__ cmask,
                           T* __restrict__ c, VLay cl, int k, const int* active) {
  const int Nc = (int)gc.nodes();
  int n, r;
  if (cl.rs == 3 && k > 1) {  // interleaved coarse: column fastest
    const int tid = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (tid >= Nc * k) return;
    n = tid / k;
    r = tid - n * k;
  } else {
    r = blockIdx.y;
    n = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (n >= Nc) return;
  }
  if (active && !active[r]) return;
  T acc[3];
  restrict_node<T>(gf, f + (long long)r * fl.rs, (int)fl.ns, gc, n, acc);
  const uint8_t m = cmask[n];
  T* C = c + (long long)n * cl.ns + (long long)r * cl.rs;
  for (int q = 0; q < 3; ++q) C[q] = live_bit(m, q) ? acc[q] : T(0);
}

// Restriction of vectors with the three components of a node adjacent (column-major, or one column): a thread per
// coarse node finds its 27 fine neighbours once and loops over a group of kXferCols columns (grid y: the groups),
// so the index arithmetic (two divisions, 27 bounds checks and addresses) is shared by the group while the lanes
// of consecutive nodes keep each column's accesses coalesced and the groups keep enough loads in flight (0.38
// against 0.58 ms for the 42-column fine-level restriction of the 256x256x48 frame).  Sums are formed in the same
// order as restrict_node, with the same results: a neighbour outside the fine grid takes a byte offset past the
// end of the column, which the buffer load reads as zero.  The prolongation keeps a thread per (node, column):
// grouping its columns made it slower (1.01 against 0.90 ms), its index work being small.
constexpr int kXferCols = 6;

template <typename T>
__global__ __launch_bounds__(256) void k_restrict_cols(Geom gf, const T* __restrict__ f, long long frs, Geom gc,
                                                       const uint8_t* __restrict__ cmask, T* __restrict__ c,
                                                       long long crs, int k, const int* active) {
  const int n = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= (int)gc.nodes()) return;
  int I, J, K;
  node_xyz(n, gc, I, J, K);
  unsigned o[27];
#pragma unroll
  for (int e = 0; e < 27; ++e) {
    const int ex = e / 9 - 1, ey = (e / 3) % 3 - 1, ez = e % 3 - 1;
    const int fi = 2 * I + ex, fj = 2 * J + ey, fk = 2 * K + ez;
    const bool in = fi >= 0 && fj >= 0 && fk >= 0 && fi < gf.nx && fj < gf.ny && fk < gf.nz;
    o[e] = in ? (unsigned)((fi * gf.ny + fj) * gf.nz + fk) * 3u * (unsigned)sizeof(T) : kOutside;
  }
  const uint8_t m = cmask[n];
  const long long colb = gf.nodes() * 3 * (long long)sizeof(T);
  const int r1 = min(k, (int)(blockIdx.y + 1) * kXferCols);
#pragma unroll 2
  for (int r = (int)blockIdx.y * kXferCols; r < r1; ++r) {
    if (active && !active[r]) continue;
    const __amdgpu_buffer_rsrc_t fr = buffer(f + (long long)r * frs, colb);
    T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int e = 0; e < 27; ++e) {
      const int ex = e / 9 - 1, ey = (e / 3) % 3 - 1, ez = e % 3 - 1;
      const T w = T((ex == 0 ? 1.0 : 0.5) * (ey == 0 ? 1.0 : 0.5) * (ez == 0 ? 1.0 : 0.5));
      T v0, v1, v2;
      load3(fr, o[e], v0, v1, v2);
      acc[0] += w * v0;
      acc[1] += w * v1;
      acc[2] += w * v2;
    }
    T* C = c + 3LL * n + (long long)r * crs;
    for (int q = 0; q < 3; ++q) C[q] = live_bit(m, q) ? acc[q] : T(0);
  }
}

template <typename T>
void restrict_full(const Geom& fine, const T* f, VLay flay, const Geom& coarse, const uint8_t* cmask, T* c,
                   VLay clay, int k, const int* active, hipStream_t s) {
  if (fine.nodes() * 3 * (long long)k > (long long)INT32_MAX) throw std::runtime_error("restrict: level too large");
  const long long nc = coarse.nodes();
  if (clay.rs == 3 && k > 1)
    hipLaunchKernelGGL(k_restrict<T>, dim3((unsigned)((nc * k + 255) / 256)), dim3(256), 0, s, fine, f, flay, coarse,
                       cmask, c, clay, k, active);
  else if (flay.ns == 3 && clay.ns == 3 && fine.nodes() * 3 * (long long)sizeof(T) < (long long)kOutside)
    hipLaunchKernelGGL(k_restrict_cols<T>, dim3((unsigned)((nc + 255) / 256), (k + kXferCols - 1) / kXferCols),
                       dim3(256), 0, s, fine, f, flay.rs,
                       coarse, cmask, c, clay.rs, k, active);
  else
    hipLaunchKernelGGL(k_restrict<T>, dim3((unsigned)((nc + 255) / 256), k), dim3(256), 0, s, fine, f, flay, coarse,
                       cmask, c, clay, k, active);
  DFMI_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_prolong_add(Geom gc, const T* __restrict__ c, VLay cl, Geom gf, const uint8_t* __restrict__ fmask,
                              T* __restrict__ f, VLay fl, int k, const int* active) {
  const int Nf = (int)gf.nodes();
  int n, r;
  if (fl.rs == 3 && k > 1) {  // interleaved fine: column fastest
    const int tid = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (tid >= Nf * k) return;
    n = tid / k;
    r = tid - n * k;
  } else {
    r = blockIdx.y;
    n = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (n >= Nf) return;
  }
  if (active && !active[r]) return;
  T acc[3];
  prolong_node<T>(gc, c + (long long)r * cl.rs, (int)cl.ns, gf, n, acc);
  const uint8_t m = fmask[n];
  T* F = f + (long long)n * fl.ns + (long long)r * fl.rs;
  // the three components are read and written unconditionally (one 3-word access each way)
  const T f0 = F[0], f1 = F[1], f2 = F[2];
  F[0] = live_bit(m, 0) ? f0 + acc[0] : f0;
  F[1] = live_bit(m, 1) ? f1 + acc[1] : f1;
  F[2] = live_bit(m, 2) ? f2 + acc[2] : f2;
}

template <typename T>
void prolong_add(const Geom& coarse, const T* c, VLay clay, const Geom& fine, const uint8_t* fmask, T* f,
                 VLay flay, int k, const int* active, hipStream_t s) {
  if (fine.nodes() * 3 * (long long)k > (long long)INT32_MAX) throw std::runtime_error("prolong: level too large");
  const long long nf = fine.nodes();
  if (flay.rs == 3 && k > 1)
    hipLaunchKernelGGL(k_prolong_add<T>, dim3((unsigned)((nf * k + 255) / 256)), dim3(256), 0, s, coarse, c, clay,
                       fine, fmask, f, flay, k, active);
  else
    hipLaunchKernelGGL(k_prolong_add<T>, dim3((unsigned)((nf + 255) / 256), k), dim3(256), 0, s, coarse, c, clay,
                       fine, fmask, f, flay, k, active);
  DFMI_HIP(hipGetLastError());
}

// Layout change of a level vector between RHS-interleaved [node][r][3] and column-major [r][node][3]
// through an LDS tile of RN nodes x all columns, so that both the reads and the writes are contiguous
// runs (the fine level is column-major, the coarse levels interleaved).
constexpr int RN = 32;
template <typename T, bool TO_COLUMNS>
__global__ __launch_bounds__(256) void k_relayout(const T* __restrict__ in, T* __restrict__ out, int N, int k,
                                                  const int* active) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);  // [RN][3k + 1]
  const int w = 3 * k, ld = w + 1;
  const int n0 = (int)xcd_block(blockIdx.x, gridDim.x) * RN;
  const int cnt = min(RN, N - n0);
  if (TO_COLUMNS) {
    for (int v = threadIdx.x; v < cnt * w; v += 256) {
      const int j = v / w, e = v - j * w;
      tile[j * ld + e] = in[(long long)n0 * w + v];
    }
    __syncthreads();
    for (int v = threadIdx.x; v < k * cnt * 3; v += 256) {
      const int r = v / (cnt * 3), rem = v - r * (cnt * 3), j = rem / 3, c = rem - j * 3;
      if (active && !active[r]) continue;
      out[((long long)r * N + n0 + j) * 3 + c] = tile[j * ld + 3 * r + c];
    }
  } else {
    for (int v = threadIdx.x; v < k * cnt * 3; v += 256) {
      const int r = v / (cnt * 3), rem = v - r * (cnt * 3), j = rem / 3, c = rem - j * 3;
      tile[j * ld + 3 * r + c] = in[((long long)r * N + n0 + j) * 3 + c];
    }
    __syncthreads();
    for (int v = threadIdx.x; v < cnt * w; v += 256) {
      const int j = v / w, e = v - j * w;
      if (active && !active[e / 3]) continue;
      out[(long long)n0 * w + v] = tile[j * ld + e];
    }
  }
}

template <typename T>
void relayout(const T* in, T* out, long long nodes, int k, bool to_columns, const int* active, hipStream_t s) {
  if (nodes * 3 * (long long)k > (long long)INT32_MAX) throw std::runtime_error("relayout: level too large");
  const unsigned blocks = (unsigned)((nodes + RN - 1) / RN);
  const size_t lds = sizeof(T) * (size_t)RN * (3 * k + 1);
  if (lds > 64 * 1024) throw std::runtime_error("relayout: too many right-hand sides for one tile");
  if (to_columns)
    hipLaunchKernelGGL((k_relayout<T, true>), dim3(blocks), dim3(256), lds, s, in, out, (int)nodes, k, active);
  else
    hipLaunchKernelGGL((k_relayout<T, false>), dim3(blocks), dim3(256), lds, s, in, out, (int)nodes, k, active);
  DFMI_HIP(hipGetLastError());
}
template void relayout<float>(const float*, float*, long long, int, bool, const int*, hipStream_t);
template void relayout<double>(const double*, double*, long long, int, bool, const int*, hipStream_t);

template void restrict_full<float>(const Geom&, const float*, VLay, const Geom&, const uint8_t*, float*, VLay, int,
                                   const int*, hipStream_t);
template void restrict_full<double>(const Geom&, const double*, VLay, const Geom&, const uint8_t*, double*, VLay, int,
                                    const int*, hipStream_t);
template void prolong_add<float>(const Geom&, const float*, VLay, const Geom&, const uint8_t*, float*, VLay, int,
                                 const int*, hipStream_t);
template void prolong_add<double>(const Geom&, const double*, VLay, const Geom&, const uint8_t*, double*, VLay, int,
                                  const int*, hipStream_t);

// ---- coarsest level: dense inverse of the live-DOF Galerkin matrix -------------------------
__global__ void k_coarsest_assemble(Geom g, const double* __restrict__ st, const int* __restrict__ live_index, int n,
                                    double* __restrict__ dense) {
  const long long N = g.nodes();
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (node >= N) return;
  const int l = (int)(node % g.nz), j = (int)((node / g.nz) % g.ny), i = (int)(node / ((long long)g.nz * g.ny));
  const int mi = i + d / 9 - 1, mj = j + (d / 3) % 3 - 1, ml = l + d % 3 - 1;
  if (mi < 0 || mj < 0 || ml < 0 || mi >= g.nx || mj >= g.ny || ml >= g.nz) return;
  const long long m = ((long long)mi * g.ny + mj) * g.nz + ml;
  for (int c = 0; c < 3; ++c) {
    const int row = live_index[node * 3 + c];
    if (row < 0) continue;
    for (int cc = 0; cc < 3; ++cc) {
      const int col = live_index[m * 3 + cc];
      if (col < 0) continue;
      dense[(long long)row * n + col] = st[((long long)d * 9 + 3 * c + cc) * N + node];
    }
  }
}

__global__ void k_symmetrize(double* a, int n) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  if (j <= i) return;
  const double v = 0.5 * (a[(long long)i * n + j] + a[(long long)j * n + i]);
  a[(long long)i * n + j] = v;
  a[(long long)j * n + i] = v;
}

void coarsest_assemble(const Geom& g, const double* st, const int* live_index, int n, double* dense,
                       hipStream_t s) {
  DFMI_HIP(hipMemsetAsync(dense, 0, sizeof(double) * (size_t)n * n, s));
  hipLaunchKernelGGL(k_coarsest_assemble, dim3((unsigned)((g.nodes() + 255) / 256), 27), dim3(256), 0, s, g, st,
                     live_index, n, dense);
  DFMI_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)(((long long)n * n + 255) / 256)), dim3(256), 0, s, dense, n);
  DFMI_HIP(hipGetLastError());
}

// Symmetric sweep operator (Goodnight 1979), blocked: sweeping the pivot block P of a symmetric
// matrix maps [A_PP A_PQ; A_QP A_QQ] to [-A_PP^-1, A_PP^-1 A_PQ; A_QP A_PP^-1, A_QQ - A_QP A_PP^-1 A_PQ];
// sweeps on disjoint blocks commute, and sweeping every pivot of an SPD matrix leaves -A^-1.
// Per block of NB pivots: copy the panel rows R = A[P, :], invert the NB x NB diagonal block in
// LDS (Gauss-Jordan, no pivoting: SPD), form W = A_PP^-1 R, then one rank-NB update of A.
constexpr int NB = 32;

__global__ void k_panel_copy(const double* __restrict__ a, int n, int k0, int b, double* __restrict__ r) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)b * n) return;
  const int p = (int)(idx / n), j = (int)(idx % n);
  r[idx] = a[(long long)(k0 + p) * n + j];
}

// Gauss-Jordan inverse of the b x b SPD block R[:, k0:k0+b] -> dinv, in one wavefront: lane j keeps column j
// of the block in registers; at pivot k the pivot column is broadcast from lane k through the scalar unit
// (readlane), so no barrier or LDS round trip separates the pivots.
__device__ __forceinline__ double lane_value(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(64) void k_block_inverse(const double* __restrict__ r, int n, int k0, int b,
                                                      double* __restrict__ dinv) {
  const int j = threadIdx.x;
  double col[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) col[i] = (i < b && j < b) ? r[(long long)i * n + k0 + j] : 0.0;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    if (k < b) {  // uniform; the loop is fully unrolled so that col stays in registers
      const double inv = 1.0 / lane_value(col[k], k);
      const double f = col[k] * inv;  // m[k][j] / d
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        if (i == k) continue;
        const double ck = lane_value(col[i], k);  // m[i][k]
        col[i] = j == k ? ck * inv : col[i] - ck * f;
      }
      col[k] = j == k ? -inv : f;
    }
  }
  // the block now holds -D^-1
  if (j < b) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < b) dinv[i * b + j] = -col[i];
  }
}

// W[p][j] = sum_q dinv[p][q] R[q][j]
__global__ void k_panel_w(const double* __restrict__ dinv, const double* __restrict__ r, int n, int b,
                          double* __restrict__ w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (j >= n) return;
  double acc = 0.0;
  for (int q = 0; q < b; ++q) acc = fma(dinv[p * b + q], r[(long long)q * n + j], acc);
  w[(long long)p * n + j] = acc;
}

// rank-b update; 64 x 64 output tile per workgroup, 4 x 4 outputs per thread
__global__ __launch_bounds__(256) void k_block_update(double* __restrict__ a, int n, int k0, int b,
                                                      const double* __restrict__ r, const double* __restrict__ w,
                                                      const double* __restrict__ dinv) {
  __shared__ double ri[NB][64];
  __shared__ double wj[NB][64];
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  for (int v = threadIdx.x; v < b * 64; v += 256) {
    const int p = v / 64, c = v % 64;
    ri[p][c] = (i0 + c < n) ? r[(long long)p * n + i0 + c] : 0.0;
    wj[p][c] = (j0 + c < n) ? w[(long long)p * n + j0 + c] : 0.0;
  }
  __syncthreads();
  const int tx = threadIdx.x % 16, tyy = threadIdx.x / 16;
  double acc[4][4] = {};
  for (int p = 0; p < b; ++p) {
    double x[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      x[u] = ri[p][tyy + 16 * u];
      y[u] = wj[p][tx + 16 * u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = fma(x[u], y[v], acc[u][v]);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + tyy + 16 * u, j = j0 + tx + 16 * v;
      if (i >= n || j >= n) continue;
      const bool pi = i >= k0 && i < k0 + b, pj = j >= k0 && j < k0 + b;
      double& dst = a[(long long)i * n + j];
      if (pi && pj) dst = -dinv[(i - k0) * b + (j - k0)];
      else if (pi) dst = w[(long long)(i - k0) * n + j];
      else if (pj) dst = w[(long long)(j - k0) * n + i];
      else dst -= acc[u][v];
    }
}

__global__ void k_negate(double* a, long long n) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) a[idx] = -a[idx];
}

void dense_sweep_inverse(double* a, int n, hipStream_t s) {
  double *r = nullptr, *w = nullptr, *dinv = nullptr;
  DFMI_HIP(hipMallocAsync((void**)&r, sizeof(double) * NB * (size_t)n, s));
  DFMI_HIP(hipMallocAsync((void**)&w, sizeof(double) * NB * (size_t)n, s));
  DFMI_HIP(hipMallocAsync((void**)&dinv, sizeof(double) * NB * NB, s));
  const unsigned tiles = (unsigned)((n + 63) / 64);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int b = std::min(NB, n - k0);
    hipLaunchKernelGGL(k_panel_copy, dim3((unsigned)(((long long)b * n + 255) / 256)), dim3(256), 0, s, a, n, k0, b,
                       r);
    hipLaunchKernelGGL(k_block_inverse, dim3(1), dim3(64), 0, s, r, n, k0, b, dinv);
    hipLaunchKernelGGL(k_panel_w, dim3((unsigned)((n + 255) / 256), b), dim3(256), 0, s, dinv, r, n, b, w);
    hipLaunchKernelGGL(k_block_update, dim3(tiles, tiles), dim3(256), 0, s, a, n, k0, b, r, w, dinv);
  }
  const unsigned b2 = (unsigned)(((long long)n * n + 255) / 256);
  hipLaunchKernelGGL(k_negate, dim3(b2), dim3(256), 0, s, a, (long long)n * n);
  DFMI_HIP(hipGetLastError());
  DFMI_HIP(hipFreeAsync(r, s));
  DFMI_HIP(hipFreeAsync(w, s));
  DFMI_HIP(hipFreeAsync(dinv, s));
}

template <typename T>
__global__ void k_gather_live(const T* __restrict__ v, VLay lay, const int* __restrict__ live_dofs, int n,
                              double* __restrict__ out, const int* active) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (idx >= n) return;
  const int d = live_dofs[idx];
  out[(long long)r * n + idx] =
      (active && !active[r]) ? 0.0 : (double)v[(d / 3) * lay.ns + (long long)r * lay.rs + d % 3];
}

// Coarsest solve out = in A^-1 (A^-1 symmetric, row-major n x n) on the FP64 matrix cores, split over the i
// range: each wavefront owns TJ 16-column tiles of A^-1 and accumulates the 16 x 16 output tiles of up to RT
// 16-row blocks of right-hand sides with v_mfma_f64_16x16x4f64 (one read of the right-hand sides serves TJ
// tiles).  In a 16-row step lane l (c = l % 16, q = l / 16) reads in[r][i0 + 4q + s] (r the lane's right-hand
// side c of the block) and A^-1[i0 + 4q + s][j] (j the lane's column c of the tile) and feeds product s with
// them, so both operands of a product meet at the same i; the lane then holds C[q + 4v][c], v = 0..3.  The next
// step's operands are loaded before the current step's products.  The slices are summed in a fixed order
// afterwards, so results are deterministic.
constexpr int DSPLIT = 32, DTJ = 2;
typedef double dense_d4 __attribute__((ext_vector_type(4)));

template <int RT>
__global__ __launch_bounds__(256) void k_dense_mfma(const double* __restrict__ in, const double* __restrict__ ainv,
                                                    int n, int k, int r0, int per, double* __restrict__ partial) {
  constexpr int TJ = DTJ;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  const int split = blockIdx.y;
  const int nb = (n + 15) / 16;
  const int b0 = split * per, b1 = min(nb, b0 + per);
  int j[TJ];
  bool jin[TJ];
#pragma unroll
  for (int u = 0; u < TJ; ++u) {
    j[u] = ((blockIdx.x * 4 + wave) * TJ + u) * 16 + c;
    jin[u] = j[u] < n;
  }
  const double* arow[RT];
  bool rin[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int r = r0 + 16 * t + c;
    rin[t] = r < k;
    arow[t] = in + (long long)(rin[t] ? r : 0) * n;
  }
  dense_d4 acc[RT][TJ];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int u = 0; u < TJ; ++u) acc[t][u] = dense_d4{0.0, 0.0, 0.0, 0.0};
  double av[RT][4], bv[TJ][4];
  auto load = [&](int blk) {
    const int i = blk * 16 + 4 * q;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const bool ok = i + s < n;
#pragma unroll
      for (int u = 0; u < TJ; ++u) bv[u][s] = ok && jin[u] ? ainv[(long long)(i + s) * n + j[u]] : 0.0;
#pragma unroll
      for (int t = 0; t < RT; ++t) av[t][s] = ok && rin[t] ? arow[t][i + s] : 0.0;
    }
  };
  if (b0 < b1) load(b0);
  for (int blk = b0; blk < b1; ++blk) {
    double ca[RT][4], cb[TJ][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
      for (int u = 0; u < TJ; ++u) cb[u][s] = bv[u][s];
#pragma unroll
      for (int t = 0; t < RT; ++t) ca[t][s] = av[t][s];
    }
    if (blk + 1 < b1) load(blk + 1);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int u = 0; u < TJ; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[t][s], cb[u][s], acc[t][u], 0, 0, 0);
  }
#pragma unroll
  for (int u = 0; u < TJ; ++u) {
    if (!jin[u]) continue;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = r0 + 16 * t + q + 4 * v;
        if (r < k) partial[((long long)split * k + r) * n + j[u]] = acc[t][u][v];
      }
  }
}

__global__ void k_sum_splits(const double* __restrict__ partial, int splits, long long kn, double* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kn) return;
  double s = 0.0;
  for (int c = 0; c < splits; ++c) s += partial[(long long)c * kn + e];
  out[e] = s;
}

int coarsest_splits() { return DSPLIT; }

template <typename T>
__global__ void k_scatter_live(const double* __restrict__ in, VLay lay, const int* __restrict__ live_dofs, int n,
                               T* __restrict__ v, const int* active) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (idx >= n || (active && !active[r])) return;
  const int d = live_dofs[idx];
  v[(d / 3) * lay.ns + (long long)r * lay.rs + d % 3] = (T)in[(long long)r * n + idx];
}

template <typename T>
void coarsest_solve(const Geom& g, const double* ainv, const int* live_dofs, int n, const T* rhs, T* out, VLay lay,
                    double* work_in, double* work_out, double* work_split, int k, const int* active, hipStream_t s) {
  DFMI_HIP(hipMemsetAsync(out, 0, sizeof(T) * (size_t)k * g.nodes() * 3, s));
  if (n == 0) return;
  const unsigned b = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_gather_live<T>, dim3(b, k), dim3(256), 0, s, rhs, lay, live_dofs, n, work_in, active);
  const int nb = (n + 15) / 16;
  const int per = (nb + DSPLIT - 1) / DSPLIT;
  const int splits = (nb + per - 1) / per;
  const dim3 grid((unsigned)((nb + 4 * DTJ - 1) / (4 * DTJ)), (unsigned)splits);
  for (int r0 = 0; r0 < k; r0 += 64) {  // up to four 16-row blocks of right-hand sides per launch
    switch (std::min(4, (k - r0 + 15) / 16)) {
      case 1: hipLaunchKernelGGL(k_dense_mfma<1>, grid, dim3(256), 0, s, work_in, ainv, n, k, r0, per, work_split); break;
      case 2: hipLaunchKernelGGL(k_dense_mfma<2>, grid, dim3(256), 0, s, work_in, ainv, n, k, r0, per, work_split); break;
      case 3: hipLaunchKernelGGL(k_dense_mfma<3>, grid, dim3(256), 0, s, work_in, ainv, n, k, r0, per, work_split); break;
      default: hipLaunchKernelGGL(k_dense_mfma<4>, grid, dim3(256), 0, s, work_in, ainv, n, k, r0, per, work_split);
    }
  }
  hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)(((long long)k * n + 255) / 256)), dim3(256), 0, s, work_split,
                     splits, (long long)k * n, work_out);
  hipLaunchKernelGGL(k_scatter_live<T>, dim3(b, k), dim3(256), 0, s, work_out, lay, live_dofs, n, out, active);
  DFMI_HIP(hipGetLastError());
}

template void coarsest_solve<float>(const Geom&, const double*, const int*, int, const float*, float*, VLay, double*,
                                    double*, double*, int, const int*, hipStream_t);
template void coarsest_solve<double>(const Geom&, const double*, const int*, int, const double*, double*, VLay,
                                     double*, double*, double*, int, const int*, hipStream_t);

}  // namespace dfmi
