// Fine-level matrix-free hex8 operator y = M A(E) M x on a regular grid.
//
// Element stiffness in the corner-Hadamard basis: with u_q (q = a + 2b + 4c the
// corner offset bits) transformed per component by the 2x2x2 Hadamard matrix H,
// Ke = H^T B H / 64 where B is block diagonal with eight 3x3 blocks (one per
// reflection-parity class p = m xor e_comp).  Per element and right-hand side
// that is 72 add/sub (forward butterfly) + 72 FMA (blocks) + 72 add/sub
// (inverse butterfly) instead of 576 FMA for the dense 24x24 product, and the
// rigid translations are annihilated exactly (their Hadamard mode is 0).
//
// Work decomposition: one workgroup owns a (TY-1) x (TZ-1) tile of node columns
// in (y, z) (z contiguous in memory, so loads coalesce) and marches along x.  Each
// step stages node plane i+1 of the operand in LDS, every thread evaluates one
// element of layer i (the tile plus a one-element halo at the low side), the 8
// corner results go through LDS, and each thread sums the four elements of layer
// i and carries the four of layer i+1 in registers.  Sums are in a fixed order,
// so results are deterministic.
#include "dfmi_internal.h"

namespace dfmi {

template <typename T>
struct Blocks {
  T b[8][9];
};

template <typename T>
__device__ __forceinline__ void butterfly(T (&u)[8][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int bit = 1; bit < 8; bit <<= 1) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (!(q & bit)) {
          T lo = u[q][c], hi = u[q | bit][c];
          u[q][c] = lo + hi;
          u[q | bit][c] = lo - hi;
        }
      }
    }
  }
}

// w = s * B v on Hadamard modes (class p holds modes p ^ e_c for components c).
template <typename T>
__device__ __forceinline__ void blocks_mul(const Blocks<T>& kb, const T (&v)[8][3], T (&w)[8][3], T s) {
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const T v0 = v[p ^ 1][0], v1 = v[p ^ 2][1], v2 = v[p ^ 4][2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T acc = kb.b[p][3 * c + 0] * v0;
      acc = fma(kb.b[p][3 * c + 1], v1, acc);
      acc = fma(kb.b[p][3 * c + 2], v2, acc);
      w[p ^ (1 << c)][c] = s * acc;
    }
  }
}

__host__ __device__ inline bool bit_live(uint8_t m, int c) { return (m >> c) & 1; }

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_fine_apply(Geom g, const T* __restrict__ moduli,
                                                     const uint8_t* __restrict__ mask, Blocks<T> kb, ApplyArgs a,
                                                     int TY, int TZ, int xchunk, int tiles_y, int tiles_z) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* U = reinterpret_cast<T*>(smem_raw);                   // [2][(TY+1)*(TZ+1)][3]
  const int plane = (TY + 1) * (TZ + 1);
  T* OUT = U + 2 * plane * 3;                               // [8][256][3]
  __shared__ double red[4];

  const int r = blockIdx.y;
  if (a.active && !a.active[r]) return;
  int tile = blockIdx.x;
  const int tzi = tile % tiles_z;
  tile /= tiles_z;
  const int tyi = tile % tiles_y;
  const int cx = tile / tiles_y;
  const int j0 = tyi * (TY - 1), l0 = tzi * (TZ - 1);
  const int i0 = cx * xchunk;
  const int i1 = min(i0 + xchunk, g.nx);
  const int t = threadIdx.x;
  const bool live = t < TY * TZ;
  const int ty = live ? t / TZ : 0, tz = live ? t % TZ : 0;
  const int ej = j0 - 1 + ty, el = l0 - 1 + tz;  // element coords of this thread == node coords of its output
  const long long nodes = g.nodes();
  const T* X = reinterpret_cast<const T*>(a.x) + (long long)r * nodes * 3;
  T* Y = reinterpret_cast<T*>(a.y) + (long long)r * nodes * 3;
  const bool own = live && ty >= 1 && tz >= 1 && ej < g.ny && el < g.nz;
  const bool elem_in = live && ej >= 0 && el >= 0 && ej < g.ey && el < g.ez;

  auto load_plane = [&](int ip, T* dst) {
    for (int v = t; v < plane * 3; v += 256) {
      const int nl = v / 3, c = v - 3 * (v / 3);
      const int ly = nl / (TZ + 1), lz = nl - ly * (TZ + 1);
      const int gj = j0 - 1 + ly, gl = l0 - 1 + lz;
      T val = T(0);
      if (ip >= 0 && ip < g.nx && gj >= 0 && gj < g.ny && gl >= 0 && gl < g.nz) {
        const long long n = ((long long)ip * g.ny + gj) * g.nz + gl;
        if (bit_live(mask[n], c)) val = X[n * 3 + c];
      }
      dst[v] = val;
    }
  };

  T carry[3] = {T(0), T(0), T(0)};
  double dot = 0.0;
  load_plane(i0 - 1, U + ((i0 - 1) & 1) * plane * 3);
  for (int i = i0 - 1; i < i1; ++i) {
    load_plane(i + 1, U + ((i + 1) & 1) * plane * 3);
    __syncthreads();
    // own operand value of node plane i, taken now: the next step's plane load reuses buffer i & 1
    // before the epilogue of this step has finished on every thread
    T xs[3] = {T(0), T(0), T(0)};
    if (own) {
      const T* src = U + (i & 1) * plane * 3 + (ty * (TZ + 1) + tz) * 3;
      xs[0] = src[0];
      xs[1] = src[1];
      xs[2] = src[2];
    }
    T y[8][3];
    if (elem_in && i >= 0 && i < g.ex) {
      const T e = moduli[((long long)i * g.ey + ej) * g.ez + el];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int qa = q & 1, qb = (q >> 1) & 1, qc = (q >> 2) & 1;
        const T* src = U + ((i + qa) & 1) * plane * 3 + ((ty + qb) * (TZ + 1) + tz + qc) * 3;
        y[q][0] = src[0];
        y[q][1] = src[1];
        y[q][2] = src[2];
      }
      butterfly(y);
      T w[8][3];
      blocks_mul(kb, y, w, e);
      butterfly(w);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        y[q][0] = w[q][0];
        y[q][1] = w[q][1];
        y[q][2] = w[q][2];
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) y[q][0] = y[q][1] = y[q][2] = T(0);
    }
    if (live) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        T* d = OUT + (q * 256 + t) * 3;
        d[0] = y[q][0];
        d[1] = y[q][1];
        d[2] = y[q][2];
      }
    }
    __syncthreads();
    if (own) {
      T s0[3] = {carry[0], carry[1], carry[2]};
      T s1[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int bc = 0; bc < 4; ++bc) {
        const int qb = bc & 1, qc = bc >> 1;
        const int src = (ty - qb) * TZ + (tz - qc);
        const T* o0 = OUT + ((2 * qb + 4 * qc) * 256 + src) * 3;
        const T* o1 = OUT + ((1 + 2 * qb + 4 * qc) * 256 + src) * 3;
        s0[0] += o0[0];
        s0[1] += o0[1];
        s0[2] += o0[2];
        s1[0] += o1[0];
        s1[1] += o1[1];
        s1[2] += o1[2];
      }
      carry[0] = s1[0];
      carry[1] = s1[1];
      carry[2] = s1[2];
      if (i >= i0) {
        const long long n = ((long long)i * g.ny + ej) * g.nz + el;
        const uint8_t mk = mask[n];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const bool on = bit_live(mk, c);
          const long long o = (long long)r * nodes * 3 + n * 3 + c;
          if (MODE == MODE_APPLY) {
            Y[n * 3 + c] = on ? s0[c] : T(0);
          } else if (MODE == MODE_DOT) {
            const T yv = on ? s0[c] : T(0);
            Y[n * 3 + c] = yv;
            dot += (double)xs[c] * (double)yv;
          } else if (MODE == MODE_RESID) {
            const T bv = reinterpret_cast<const T*>(a.b)[o];
            Y[n * 3 + c] = on ? bv - s0[c] : T(0);
          } else if (MODE == MODE_CHEB) {
            const T bv = reinterpret_cast<const T*>(a.b)[o];
            const T xpv = a.xp ? reinterpret_cast<const T*>(a.xp)[o] : T(0);
            const T dv = reinterpret_cast<const T*>(a.dinv)[n * 3 + c];
            const T c1 = (T)a.c1, c2 = (T)a.c2;
            const T xv = xs[c];
            Y[n * 3 + c] = on ? xv + c1 * (xv - xpv) + c2 * dv * (bv - s0[c]) : T(0);
          }
        }
      }
    }
  }
  if (MODE == MODE_DOT) {
    // wavefront reduction (64 lanes) then across the 4 wavefronts of the block
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
    if ((t & 63) == 0) red[t >> 6] = dot;
    __syncthreads();
    if (t == 0) a.partial[(long long)r * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
  }
}

struct Tiling {
  int TY, TZ, tiles_y, tiles_z, xchunk, chunks;
};

static Tiling choose_tiling(const Geom& g, int k) {
  Tiling best{};
  double best_cost = 1e300;
  for (int TZ = 8; TZ <= 64; ++TZ) {
    int TY = 256 / TZ;
    if (TY < 3) continue;
    const int ty = (g.ny + TY - 2) / (TY - 1), tz = (g.nz + TZ - 2) / (TZ - 1);
    const int waves = (TY * TZ + 63) / 64;
    double cost = (double)ty * tz * waves * (1.0 + 0.15 * (TZ < 16) + 0.3 * (TZ < 12));
    if (cost < best_cost) {
      best_cost = cost;
      best = Tiling{TY, TZ, ty, tz, 0, 0};
    }
  }
  const long long tiles = (long long)best.tiles_y * best.tiles_z * (k > 0 ? k : 1);
  const long long target = 4096;
  int chunks = (int)((target + tiles - 1) / tiles);
  chunks = std::max(1, std::min(chunks, (g.nx + 7) / 8));
  best.xchunk = (g.nx + chunks - 1) / chunks;
  best.chunks = (g.nx + best.xchunk - 1) / best.xchunk;
  return best;
}

int fine_blocks(const Geom& g) {
  // upper bound for the partial-sum buffer (tiling for k = 1 has the most chunks)
  Tiling tl = choose_tiling(g, 1);
  return tl.tiles_y * tl.tiles_z * tl.chunks;
}

template <typename T>
void fine_apply(const Geom& g, const T* moduli, const uint8_t* mask, const KeBlocks& kbd, int mode,
                const ApplyArgs& a, hipStream_t s, int* nblocks_out) {
  Tiling tl = choose_tiling(g, mode == MODE_DOT ? 1 : a.k);
  Blocks<T> kb;
  for (int p = 0; p < 8; ++p)
    for (int e = 0; e < 9; ++e) kb.b[p][e] = (T)kbd.b[p][e / 3][e % 3];
  dim3 grid(tl.tiles_y * tl.tiles_z * tl.chunks, a.k);
  const size_t lds = (size_t)(2 * (tl.TY + 1) * (tl.TZ + 1) * 3 + 8 * 256 * 3) * sizeof(T);
  if (nblocks_out) *nblocks_out = grid.x;
#define DFMI_LAUNCH(M)                                                                                   \
  hipLaunchKernelGGL((k_fine_apply<T, M>), grid, dim3(256), lds, s, g, moduli, mask, kb, a, tl.TY, tl.TZ, \
                     tl.xchunk, tl.tiles_y, tl.tiles_z)
  switch (mode) {
    case MODE_APPLY: DFMI_LAUNCH(MODE_APPLY); break;
    case MODE_RESID: DFMI_LAUNCH(MODE_RESID); break;
    case MODE_CHEB: DFMI_LAUNCH(MODE_CHEB); break;
    case MODE_DOT: DFMI_LAUNCH(MODE_DOT); break;
    default: throw std::runtime_error("fine_apply: bad mode");
  }
#undef DFMI_LAUNCH
  DFMI_HIP(hipGetLastError());
}

template void fine_apply<float>(const Geom&, const float*, const uint8_t*, const KeBlocks&, int, const ApplyArgs&,
                                hipStream_t, int*);
template void fine_apply<double>(const Geom&, const double*, const uint8_t*, const KeBlocks&, int,
                                 const ApplyArgs&, hipStream_t, int*);

// ---- per-element energy u_e^T Ke u_e over the unpadded elements (gathered from global memory) ----
__global__ void k_element_energy(Geom pad, int ux, int uy, int uz, const double* __restrict__ u, Blocks<double> kb,
                                 double* __restrict__ energy) {
  const long long ne = (long long)ux * uy * uz;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (e >= ne) return;
  const int el = (int)(e % uz);
  const int ej = (int)((e / uz) % uy);
  const int ei = (int)(e / ((long long)uz * uy));
  const double* X = u + (long long)r * pad.nodes() * 3;
  double v[8][3];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const long long n = ((long long)(ei + (q & 1)) * pad.ny + ej + ((q >> 1) & 1)) * pad.nz + el + ((q >> 2) & 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[q][c] = X[n * 3 + c];
  }
  butterfly(v);
  double w[8][3];
  blocks_mul(kb, v, w, 1.0);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc = fma(v[q][c], w[q][c], acc);
  // v = H u and the blocks carry the 1/64: v . w = u^T H^T B H u / 64 = u^T Ke u
  energy[(long long)r * ne + e] = acc;
}

void fine_element_energy(const Geom& pad, int ux, int uy, int uz, const double* u, const KeBlocks& kbd,
                         double* energy, int k, hipStream_t s) {
  Blocks<double> kb;
  for (int p = 0; p < 8; ++p)
    for (int e = 0; e < 9; ++e) kb.b[p][e] = kbd.b[p][e / 3][e % 3];
  const long long ne = (long long)ux * uy * uz;
  dim3 grid((unsigned)((ne + 255) / 256), k);
  hipLaunchKernelGGL(k_element_energy, grid, dim3(256), 0, s, pad, ux, uy, uz, u, kb, energy);
  DFMI_HIP(hipGetLastError());
}

// ---- masks, diagonal and the assembled fine stencil (setup only) -----------------------
// mask bit c of node n: some element touching n has E > 0 and DOF 3n+c is not fixed
__global__ void k_active_mask(Geom g, const double* __restrict__ moduli, const uint8_t* __restrict__ fixed,
                              uint8_t* __restrict__ mask, int vx, int vy, int vz) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nodes()) return;
  const int l = (int)(n % g.nz), j = (int)((n / g.nz) % g.ny), i = (int)(n / ((long long)g.nz * g.ny));
  bool act = false;
  for (int q = 0; q < 8; ++q) {
    const int ei = i - (q & 1), ej = j - ((q >> 1) & 1), el = l - ((q >> 2) & 1);
    if (ei < 0 || ej < 0 || el < 0 || ei >= g.ex || ej >= g.ey || el >= g.ez) continue;
    if (moduli[((long long)ei * g.ey + ej) * g.ez + el] > 0) act = true;
  }
  uint8_t m = 0;
  if (act) {
    // fixed flags are indexed over the (vx, vy, vz) node grid; padding nodes are never fixed
    const bool inside = i < vx && j < vy && l < vz;
    const long long u = ((long long)i * vy + j) * vz + l;
    for (int c = 0; c < 3; ++c)
      if (!fixed || !inside || !fixed[u * 3 + c]) m |= (uint8_t)(1 << c);
  }
  mask[n] = m;
}

void active_mask(const Geom& g, const double* moduli, const uint8_t* fixed, uint8_t* mask, hipStream_t s) {
  hipLaunchKernelGGL(k_active_mask, dim3((unsigned)((g.nodes() + 255) / 256)), dim3(256), 0, s, g, moduli, fixed,
                     mask, g.nx, g.ny, g.nz);
  DFMI_HIP(hipGetLastError());
}

void active_mask_flat(const Geom& pad, int ux, int uy, int uz, const double* moduli, const uint8_t* fixed,
                      uint8_t* mask, hipStream_t s) {
  hipLaunchKernelGGL(k_active_mask, dim3((unsigned)((pad.nodes() + 255) / 256)), dim3(256), 0, s, pad, moduli, fixed,
                     mask, ux + 1, uy + 1, uz + 1);
  DFMI_HIP(hipGetLastError());
}

// Ke in corner-bit order: ke[(3*q + c) * 24 + 3*q' + c']
__global__ void k_fine_diagonal(Geom g, const double* __restrict__ moduli, const uint8_t* __restrict__ mask,
                                const double* __restrict__ kediag, double* __restrict__ diag) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nodes()) return;
  const int l = (int)(n % g.nz), j = (int)((n / g.nz) % g.ny), i = (int)(n / ((long long)g.nz * g.ny));
  double d[3] = {0, 0, 0};
  for (int q = 0; q < 8; ++q) {
    const int ei = i - (q & 1), ej = j - ((q >> 1) & 1), el = l - ((q >> 2) & 1);
    if (ei < 0 || ej < 0 || el < 0 || ei >= g.ex || ej >= g.ey || el >= g.ez) continue;
    const double e = moduli[((long long)ei * g.ey + ej) * g.ez + el];
    for (int c = 0; c < 3; ++c) d[c] += e * kediag[3 * q + c];
  }
  const uint8_t m = mask[n];
  for (int c = 0; c < 3; ++c) diag[n * 3 + c] = bit_live(m, c) ? d[c] : 0.0;
}

void fine_diagonal(const Geom& g, const double* moduli, const uint8_t* mask, const double* ke_diag, double* diag,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_fine_diagonal, dim3((unsigned)((g.nodes() + 255) / 256)), dim3(256), 0, s, g, moduli, mask,
                     ke_diag, diag);
  DFMI_HIP(hipGetLastError());
}

// Assembled masked 27-point block stencil of the fine operator:
// st[(d*9 + 3*c + c') * nodes + n] = M_nc M_mc' sum_{e ∋ n,m} E_e Ke[(q_n,c),(q_m,c')], m = n + off(d)
// with off(d) = (d/9 - 1, (d/3)%3 - 1, d%3 - 1).
__global__ void k_fine_stencil(Geom g, const double* __restrict__ moduli, const uint8_t* __restrict__ mask,
                               const double* __restrict__ ke, double* __restrict__ st) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (n >= g.nodes()) return;
  const long long N = g.nodes();
  const int l = (int)(n % g.nz), j = (int)((n / g.nz) % g.ny), i = (int)(n / ((long long)g.nz * g.ny));
  const int dx = d / 9 - 1, dy = (d / 3) % 3 - 1, dz = d % 3 - 1;
  const int mi = i + dx, mj = j + dy, ml = l + dz;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (mi >= 0 && mj >= 0 && ml >= 0 && mi < g.nx && mj < g.ny && ml < g.nz) {
    for (int q = 0; q < 8; ++q) {
      const int qa = q & 1, qb = (q >> 1) & 1, qc = (q >> 2) & 1;
      const int ra = qa + dx, rb = qb + dy, rc = qc + dz;
      if (ra < 0 || ra > 1 || rb < 0 || rb > 1 || rc < 0 || rc > 1) continue;
      const int ei = i - qa, ej = j - qb, el = l - qc;
      if (ei < 0 || ej < 0 || el < 0 || ei >= g.ex || ej >= g.ey || el >= g.ez) continue;
      const double e = moduli[((long long)ei * g.ey + ej) * g.ez + el];
      if (e == 0.0) continue;
      const int qm = ra + 2 * rb + 4 * rc;
      for (int c = 0; c < 3; ++c)
        for (int cc = 0; cc < 3; ++cc) acc[3 * c + cc] += e * ke[(3 * q + c) * 24 + 3 * qm + cc];
    }
    const uint8_t mn = mask[n], mm = mask[((long long)mi * g.ny + mj) * g.nz + ml];
    for (int c = 0; c < 3; ++c)
      for (int cc = 0; cc < 3; ++cc)
        if (!bit_live(mn, c) || !bit_live(mm, cc)) acc[3 * c + cc] = 0.0;
  }
  for (int e = 0; e < 9; ++e) st[((long long)d * 9 + e) * N + n] = acc[e];
}

void fine_stencil(const Geom& g, const double* moduli, const uint8_t* mask, const double* ke, double* stencil,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_fine_stencil, dim3((unsigned)((g.nodes() + 255) / 256), 27), dim3(256), 0, s, g, moduli, mask,
                     ke, stencil);
  DFMI_HIP(hipGetLastError());
}

}  // namespace dfmi
