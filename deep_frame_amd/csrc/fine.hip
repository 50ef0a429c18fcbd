// Fine-level matrix-free hex8 operator y = M A(E) M x on a regular grid.
//
// Element stiffness in the corner-Hadamard basis: with u_q (q = a + 2b + 4c the
// corner offset bits) transformed per component by the 2x2x2 Hadamard matrix H,
// Ke = H^T B H / 64 where B is block diagonal with eight 3x3 blocks (one per
// reflection-parity class p = m xor e_comp).  Per element and right-hand side
// that is 72 add/sub (forward butterfly) + 45 FMA (blocks) + 72 add/sub
// (inverse butterfly) instead of 576 FMA for the dense 24x24 product, and the
// rigid translations are annihilated exactly (their Hadamard mode is 0).
//
// Work decomposition: one workgroup owns a (TY-1) x (TZ-1) tile of node columns
// in (y, z) (z contiguous in memory, so loads coalesce) and marches along x.  Each
// step stages node plane i+1 of the operand in LDS and every thread evaluates one
// element of layer i (the tile plus a one-element halo at the low side).  H is
// factored into a (y, z) face transform and an x stage: the face transform of
// plane i+1 is reused by the next element along x (carried in registers), and the
// inverse x stage leaves a plane-i+1 face that is added to the next element's
// plane-i face before one inverse face transform (48 add/sub forward where the face
// is carried, 60 back, instead of 72 each way).  The three off-corner results
// of the plane-i face go through LDS and each thread sums the four faces of its
// node.  Sums are in a fixed order, so results are deterministic.
#include <cmath>
#include <stdexcept>

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

// 4-point Hadamard transform over the (y, z) corners of an element face: f[m], m = dy + 2 dz.
template <typename T>
__device__ __forceinline__ void face_butterfly(T (&f)[4][3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T a = f[0][c] + f[1][c], b = f[0][c] - f[1][c], d = f[2][c] + f[3][c], e = f[2][c] - f[3][c];
    f[0][c] = a + d;
    f[2][c] = a - d;
    f[1][c] = b + e;
    f[3][c] = b - e;
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

// The eight parity blocks of an axis-aligned brick (isotropic or orthotropic, any spacing) share one
// structural pattern: classes 0 and 7 are full symmetric 3x3 blocks, classes 1, 2, 4 are 2x2 blocks
// (their third row and column belong to a rigid translation and vanish), classes 3, 5, 6 are a 2x2
// block plus one diagonal entry.  33 distinct constants, 45 FMA per element and right-hand side.
template <typename T>
struct KeSym {
  T k[33];
};

template <typename T>
__device__ __forceinline__ void sym_mul(const KeSym<T>& s, const T (&v)[8][3], T (&w)[8][3]) {
  {  // class 0: (c0, m1) (c1, m2) (c2, m4)
    const T x = v[1][0], y = v[2][1], z = v[4][2];
    w[1][0] = fma(s.k[0], x, fma(s.k[1], y, s.k[2] * z));
    w[2][1] = fma(s.k[1], x, fma(s.k[3], y, s.k[4] * z));
    w[4][2] = fma(s.k[2], x, fma(s.k[4], y, s.k[5] * z));
  }
  {  // class 7: (c0, m6) (c1, m5) (c2, m3)
    const T x = v[6][0], y = v[5][1], z = v[3][2];
    w[6][0] = fma(s.k[6], x, fma(s.k[7], y, s.k[8] * z));
    w[5][1] = fma(s.k[7], x, fma(s.k[9], y, s.k[10] * z));
    w[3][2] = fma(s.k[8], x, fma(s.k[10], y, s.k[11] * z));
  }
  {  // class 1: (c1, m3) (c2, m5)
    const T y = v[3][1], z = v[5][2];
    w[3][1] = fma(s.k[12], y, s.k[13] * z);
    w[5][2] = fma(s.k[13], y, s.k[14] * z);
  }
  {  // class 2: (c0, m3) (c2, m6)
    const T x = v[3][0], z = v[6][2];
    w[3][0] = fma(s.k[15], x, s.k[16] * z);
    w[6][2] = fma(s.k[16], x, s.k[17] * z);
  }
  {  // class 4: (c0, m5) (c1, m6)
    const T x = v[5][0], y = v[6][1];
    w[5][0] = fma(s.k[18], x, s.k[19] * y);
    w[6][1] = fma(s.k[19], x, s.k[20] * y);
  }
  {  // class 3: (c0, m2) (c1, m1) block, (c2, m7) diagonal
    const T x = v[2][0], y = v[1][1];
    w[2][0] = fma(s.k[21], x, s.k[22] * y);
    w[1][1] = fma(s.k[22], x, s.k[23] * y);
    w[7][2] = s.k[24] * v[7][2];
  }
  {  // class 5: (c0, m4) (c2, m1) block, (c1, m7) diagonal
    const T x = v[4][0], z = v[1][2];
    w[4][0] = fma(s.k[25], x, s.k[26] * z);
    w[1][2] = fma(s.k[26], x, s.k[27] * z);
    w[7][1] = s.k[28] * v[7][1];
  }
  {  // class 6: (c1, m4) (c2, m2) block, (c0, m7) diagonal
    const T y = v[4][1], z = v[2][2];
    w[4][1] = fma(s.k[29], y, s.k[30] * z);
    w[2][2] = fma(s.k[30], y, s.k[31] * z);
    w[7][0] = s.k[32] * v[7][0];
  }
  w[0][0] = w[0][1] = w[0][2] = T(0);
}

// Threads per workgroup: one element column per thread.
template <typename T>
struct FineCfg {
  static constexpr int NT = sizeof(T) == 4 ? 512 : 256;
  static constexpr int MAXNP = NT + 136;  // (TY+1)(TZ+1) = TY*TZ + TY + TZ + 1 with TZ <= 128, TY >= 4
};

// One workgroup owns a (TY-1) x (TZ-1) tile of node columns (TY * TZ <= NT element columns including
// a one-element halo at the low side) and marches along x.  The FP32 DOT
// mode, which forms the fused PCG direction, is held to 80 VGPRs (six waves per SIMD, three workgroups per
// CU).  PRE: the operand is known to be zero on the DOFs the mask cuts (every vector the solver itself passes),
// so the staging loads skip the mask and read through 32-bit buffer offsets; slots outside the grid read zeros.
// KMODE: an ApplyMode, or kChebPartial for the FP32 Chebyshev step that also sums b.y per workgroup (a separate
// instantiation, so that the other Chebyshev steps carry no dot-product work).
// kResidWork: the FP64 residual written in FP32 (times a per-column factor) with partial sums of |r|^2, the
// restart of the FP32 recurrence after a reliable update.
constexpr int kChebPartial = 100, kResidWork = 102;
template <typename T, int KMODE, bool PRE>
__global__ __launch_bounds__(FineCfg<T>::NT)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && KMODE != MODE_MASS ? 6 : 1))) void k_fine_apply(
    Geom g, const T* __restrict__ moduli, const uint8_t* __restrict__ mask, KeSym<T> ks, ApplyArgs a, int TY, int TZ,
    int xchunk, int tiles_y, int tiles_z) {
  constexpr int MODE = KMODE == kChebPartial ? MODE_CHEB : (KMODE == kResidWork ? MODE_RESID : KMODE);
  constexpr bool kChebDot = KMODE == kChebPartial;
  constexpr bool kWork = KMODE == kResidWork;
  constexpr int NT = FineCfg<T>::NT;
  constexpr int MAXNP = FineCfg<T>::MAXNP;
  // the FP32 PCG's fusions (direction formed inside DOT, r.z partials from the last CHEB step)
  constexpr bool kFusedPcg = sizeof(T) == sizeof(float);
  __shared__ T U[2][3 * MAXNP];  // node planes i and i+1 of the operand (masked)
  __shared__ T OUT[9 * NT];      // face corner results m = 1..3 (dy, dz, dy dz) at plane i
  // LDS layouts: FP32 node-major (a node's three components adjacent: one address per node and fewer LDS
  // instructions; a 12-byte lane stride is conflict-free), FP64 component planes (8-byte words at a 24-byte
  // lane stride would conflict)
  constexpr bool kNodeMajor = sizeof(T) == sizeof(float);
  auto u_at = [&](int buf, int c, int v) -> T& { return kNodeMajor ? U[buf][3 * v + c] : U[buf][c * MAXNP + v]; };
  auto out_at = [&](int m, int c, int tt) -> T& { return kNodeMajor ? OUT[9 * tt + 3 * m + c] : OUT[(3 * m + c) * NT + tt]; };
  __shared__ double red[NT / 64];

  const int r = blockIdx.y;
  if (a.active && !a.active[r]) return;
  int tile = (int)xcd_block(blockIdx.x, gridDim.x);
  const int tzi = tile % tiles_z;
  tile /= tiles_z;
  const int tyi = tile % tiles_y;
  const int cx = tile / tiles_y;
  const int j0 = tyi * (TY - 1), l0 = tzi * (TZ - 1);
  const int i0 = cx * xchunk;
  const int i1 = min(i0 + xchunk, g.nx);
  const int t = threadIdx.x;
  const int NP = (TY + 1) * (TZ + 1);
  const bool live = t < TY * TZ;
  const int ty = live ? t / TZ : 0, tz = live ? t - (t / TZ) * TZ : 0;
  const int ej = j0 - 1 + ty, el = l0 - 1 + tz;  // element of this thread; also the node it outputs
  const int plane_nodes = g.ny * g.nz;
  const long long col = (long long)g.nodes() * 3;
  const T* X = reinterpret_cast<const T*>(a.x) + (long long)r * col;
  T* Y = reinterpret_cast<T*>(a.y) + (long long)r * col;
  const bool own = live && ty >= 1 && tz >= 1 && ej < g.ny && el < g.nz;
  const bool elem_in = live && ej >= 0 && el >= 0 && ej < g.ey && el < g.ez;
  const int self = ty * (TZ + 1) + tz;  // this thread's node in the staged plane

  // plane staging: node slots v = t and v = t + NT of the (TY+1) x (TZ+1) tile (offsets precomputed)
  int so[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int v = t + k * NT;
    so[k] = -2;  // -2: no slot, -1: slot outside the grid
    if (v < NP) {
      const int ly = v / (TZ + 1), lz = v - ly * (TZ + 1);
      const int gj = j0 - 1 + ly, gl = l0 - 1 + lz;
      so[k] = (gj >= 0 && gl >= 0 && gj < g.ny && gl < g.nz) ? gj * g.nz + gl : -1;
    }
  }
  T nv[2][3];
  // PRE staging: byte offset of each slot within a column at plane 0 (kOutside: no node), buffers of one column
  unsigned sb[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) sb[k] = so[k] >= 0 ? (unsigned)so[k] * 3u * (unsigned)sizeof(T) : kOutside;
  const unsigned plane_b = (unsigned)plane_nodes * 3u * (unsigned)sizeof(T);
  const long long col_b = col * (long long)sizeof(T);
  const __amdgpu_buffer_rsrc_t xr = buffer(X, col_b);
  const __amdgpu_buffer_rsrc_t br = buffer(a.b ? reinterpret_cast<const T*>(a.b) + (long long)r * col : X, col_b);
  const __amdgpu_buffer_rsrc_t dr = buffer(a.dinv ? a.dinv : (const void*)X, col_b);
  const __amdgpu_buffer_rsrc_t xpr = buffer(a.xp ? reinterpret_cast<const T*>(a.xp) + (long long)r * col : X, col_b);
  const __amdgpu_buffer_rsrc_t yr = buffer(Y, col_b);
  const __amdgpu_buffer_rsrc_t pr = buffer(a.pout ? reinterpret_cast<const T*>(a.pout) + (long long)r * col : X, col_b);
  const __amdgpu_buffer_rsrc_t mr = buffer(mask, g.nodes());
  const __amdgpu_buffer_rsrc_t wr =
      buffer(kWork ? reinterpret_cast<const float*>(a.pout) + (long long)r * col : reinterpret_cast<const float*>(X),
             col * (long long)sizeof(float));
  auto fetch_pre = [&](int ip) {
    const bool in = ip >= 0 && ip < g.nx;  // uniform
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const unsigned off = in ? sb[k] + (unsigned)ip * plane_b : kOutside;
      T v0, v1, v2;
      if (MODE == MODE_CHEB0) {  // x1 = c0 Dinv b
        T b0, b1, b2, d0, d1, d2;
        load3(br, off, b0, b1, b2);
        load3(dr, off, d0, d1, d2);
        const T c0 = (T)a.c0;
        v0 = c0 * d0 * b0;
        v1 = c0 * d1 * b1;
        v2 = c0 * d2 * b2;
      } else if (MODE == MODE_DOT && kFusedPcg && a.pout) {  // fused PCG direction: p = z + beta p_old
        load3(br, off, v0, v1, v2);
        if (a.beta) {
          const T bt = (T)a.beta[r];
          T p0, p1, p2;
          load3(xr, off, p0, p1, p2);
          v0 += bt * p0;
          v1 += bt * p1;
          v2 += bt * p2;
        }
      } else {
        load3(xr, off, v0, v1, v2);
      }
      nv[k][0] = v0;
      nv[k][1] = v1;
      nv[k][2] = v2;
    }
  };
  // branch-free staging: slots outside the grid read node 0 and are zeroed by their mask
  auto fetch_masked = [&](int ip) {
    const bool in = ip >= 0 && ip < g.nx;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      {
        const bool ok = in && so[k] >= 0;
        const int n = ok ? ip * plane_nodes + so[k] : 0;
        const uint8_t mr = mask[n];
        const uint8_t m = ok ? mr : 0;
        T v0, v1, v2;
        if (MODE == MODE_CHEB0) {  // x1 = c0 Dinv b
          const T* bs = reinterpret_cast<const T*>(a.b) + (long long)r * col + 3LL * n;
          const T* ds = reinterpret_cast<const T*>(a.dinv) + 3LL * n;
          const T c0 = (T)a.c0;
          v0 = c0 * ds[0] * bs[0];
          v1 = c0 * ds[1] * bs[1];
          v2 = c0 * ds[2] * bs[2];
        } else if (MODE == MODE_DOT && kFusedPcg && a.pout) {  // fused PCG direction: p = z + beta p_old
          const T* zs = reinterpret_cast<const T*>(a.b) + (long long)r * col + 3LL * n;
          v0 = zs[0];
          v1 = zs[1];
          v2 = zs[2];
          if (a.beta) {
            const T bt = (T)a.beta[r];
            const T* src = X + 3LL * n;
            v0 += bt * src[0];
            v1 += bt * src[1];
            v2 += bt * src[2];
          }
        } else {
          const T* src = X + 3LL * n;
          v0 = src[0];
          v1 = src[1];
          v2 = src[2];
        }
        nv[k][0] = (m & 1) ? v0 : T(0);
        nv[k][1] = (m & 2) ? v1 : T(0);
        nv[k][2] = (m & 4) ? v2 : T(0);
      }
    }
  };
  auto fetch = [&](int ip) {
    if constexpr (PRE)
      fetch_pre(ip);
    else
      fetch_masked(ip);
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (so[k] != -2) {
        const int v = t + k * NT;
        u_at(buf, 0, v) = nv[k][0];
        u_at(buf, 1, v) = nv[k][1];
        u_at(buf, 2, v) = nv[k][2];
      }
    }
  };
  auto modulus = [&](int i) -> T {
    const bool ok = elem_in && i >= 0 && i < g.ex;
    const T e = moduli[ok ? (i * g.ey + ej) * g.ez + el : 0];
    return ok ? e : T(0);
  };

  // The element's 8-point Hadamard transform factors into a (y, z) face transform and an x stage.  The
  // face transform of plane i+1 is the next element's plane-i transform, so it is computed once and
  // carried (fc); on the way back the x stage splits w into its plane-i and plane-i+1 faces, and the
  // plane-i+1 face (gc) is added to the next element's plane-i face before one inverse face transform.
  T fc[4][3], gc[4][3];
#pragma unroll
  for (int m = 0; m < 4; ++m) fc[m][0] = fc[m][1] = fc[m][2] = gc[m][0] = gc[m][1] = gc[m][2] = T(0);
  double dot = 0.0;
  fetch(i0 - 1);
  stage((i0 - 1) & 1);
  fetch(i0);
  T e_cur = modulus(i0 - 1);
  uint8_t mk_next = 0;
  for (int i = i0 - 1; i < i1; ++i) {
    stage((i + 1) & 1);
    if (i + 2 <= i1) fetch(i + 2);
    const T e_next = modulus(i + 1);
    const int nn = i * plane_nodes + ej * g.nz + el;
    const bool out = own && i >= i0;
    const int nnc = out ? nn : 0;  // loads below are unconditional (node 0 when this thread has no output)
    uint8_t mk;
    if constexpr (PRE) {
      mk = (uint8_t)__builtin_amdgcn_raw_buffer_load_b8(mr, out ? (unsigned)nn : kOutside, 0, 0);
    } else {
      const uint8_t mkr = mask[nnc];
      mk = out ? mkr : 0;
    }
    // epilogue operands of node (i, ej, el), requested early (PRE: 32-bit buffer offsets, zeros without a node)
    T bv[3] = {T(0), T(0), T(0)}, xpv[3] = {T(0), T(0), T(0)}, dv[3] = {T(0), T(0), T(0)};
    const unsigned eoff = out ? (unsigned)nn * 3u * (unsigned)sizeof(T) : kOutside;
    if constexpr (PRE) {
      if (MODE == MODE_RESID || MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_CHEB1)
        load3(br, eoff, bv[0], bv[1], bv[2]);
      if (MODE == MODE_CHEB && a.xp) load3(xpr, eoff, xpv[0], xpv[1], xpv[2]);
      if (MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_CHEB1) load3(dr, eoff, dv[0], dv[1], dv[2]);
    } else {
      if (MODE == MODE_RESID || MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_CHEB1) {
        const T* bp = reinterpret_cast<const T*>(a.b) + (long long)r * col + 3LL * nnc;
        bv[0] = bp[0];
        bv[1] = bp[1];
        bv[2] = bp[2];
      }
      if (MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_CHEB1) {
        if (MODE == MODE_CHEB && a.xp) {
          const T* pp = reinterpret_cast<const T*>(a.xp) + (long long)r * col + 3LL * nnc;
          xpv[0] = pp[0];
          xpv[1] = pp[1];
          xpv[2] = pp[2];
        }
        const T* dp = reinterpret_cast<const T*>(a.dinv) + 3LL * nnc;
        dv[0] = dp[0];
        dv[1] = dp[1];
        dv[2] = dp[2];
      }
    }
    __syncthreads();
    const int b0 = i & 1, b1 = (i + 1) & 1;
    T xs[3];
    xs[0] = u_at(b0, 0, self);
    xs[1] = u_at(b0, 1, self);
    xs[2] = u_at(b0, 2, self);
    // modes with epilogue operands in registers recompute the plane-i face instead of carrying it
    constexpr bool carry_face = !(MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_CHEB1);
    if (!carry_face || i == i0 - 1) {  // face transform of plane i
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int v = self + (m & 1) * (TZ + 1) + (m >> 1);
        fc[m][0] = u_at(b0, 0, v);
        fc[m][1] = u_at(b0, 1, v);
        fc[m][2] = u_at(b0, 2, v);
      }
      face_butterfly(fc);
    }
    T g[4][3];
    {  // every element is evaluated (E = 0 gives zero contributions)
      T fn[4][3];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int v = self + (m & 1) * (TZ + 1) + (m >> 1);
        fn[m][0] = u_at(b1, 0, v);
        fn[m][1] = u_at(b1, 1, v);
        fn[m][2] = u_at(b1, 2, v);
      }
      face_butterfly(fn);
      T y[8][3];  // mode q = x + 2 m
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          y[2 * m][c] = fc[m][c] + fn[m][c];
          y[2 * m + 1][c] = fc[m][c] - fn[m][c];
          if (carry_face) fc[m][c] = fn[m][c];
        }
      T w[8][3];
      if constexpr (MODE == MODE_MASS) {  // the mass is diagonal in the Hadamard basis, every mode included
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const T m = e_cur * (T)a.mdiag[q];
          w[q][0] = m * y[q][0];
          w[q][1] = m * y[q][1];
          w[q][2] = m * y[q][2];
        }
      } else {
        sym_mul(ks, y, w);  // unscaled: the modulus enters in the x stage below
      }
      // x stage of the inverse transform: plane-i face (this element, times its modulus: 12 products instead
      // of 21 on the modes) plus the previous element's plane-i face, then one inverse face transform gives the
      // four corner sums at plane i
      const T es = MODE == MODE_MASS ? T(1) : e_cur;  // the mass path scaled its modes already
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          g[m][c] = fma(es, w[2 * m][c] + w[2 * m + 1][c], gc[m][c]);
          gc[m][c] = es * (w[2 * m][c] - w[2 * m + 1][c]);
        }
      face_butterfly(g);
    }
#pragma unroll
    for (int m = 1; m < 4; ++m) {
      out_at(m - 1, 0, t) = g[m][0];
      out_at(m - 1, 1, t) = g[m][1];
      out_at(m - 1, 2, t) = g[m][2];
    }
    __syncthreads();
    if (own) {
      const int tb = t - TZ, tc = t - 1, td = t - TZ - 1;  // elements (ty-1,tz), (ty,tz-1), (ty-1,tz-1)
      T s0[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) s0[c] = g[0][c] + out_at(0, c, tb) + out_at(1, c, tc) + out_at(2, c, td);
      if (PRE || out) {  // PRE: a thread without an output node stores to kOutside, which is dropped
        T yv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const bool on = bit_live(mk, c);
          if (MODE == MODE_APPLY || MODE == MODE_DOT) {
            yv[c] = on ? s0[c] : T(0);
          } else if (MODE == MODE_RESID) {
            yv[c] = on ? bv[c] - s0[c] : T(0);
          } else if (MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_CHEB1) {
            const T c1 = (T)a.c1, c2 = (T)a.c2;
            const T prev = MODE == MODE_CHEB1 ? (T)a.c0 * dv[c] * bv[c] : xpv[c];
            yv[c] = on ? xs[c] + c1 * (xs[c] - prev) + c2 * dv[c] * (bv[c] - s0[c]) : T(0);
          } else if (MODE == MODE_MASS) {
            const T lump = a.dinv ? reinterpret_cast<const T*>(a.dinv)[3LL * nnc + c] : T(0);
            yv[c] = on ? s0[c] + lump * xs[c] : T(0);
          }
        }
        // partial sums: x.y (DOT), b.y (FP32 CHEB with partials); yv is zero without an output node
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (MODE == MODE_DOT) dot += (double)xs[c] * (double)yv[c];
          if (kChebDot) dot += (double)bv[c] * (double)yv[c];
          if (kWork) dot += (double)yv[c] * (double)yv[c];
        }
        if constexpr (kWork) {  // the FP32 residual only (PRE: offsets in floats, dropped without a node)
          const T f = (T)a.beta[r];
          store3(wr, out ? (unsigned)nn * 3u * (unsigned)sizeof(float) : kOutside, (float)(f * yv[0]),
                 (float)(f * yv[1]), (float)(f * yv[2]));
        } else if constexpr (PRE) {
          store3(yr, eoff, yv[0], yv[1], yv[2]);
          if (MODE == MODE_DOT && kFusedPcg && a.pout) store3(pr, eoff, xs[0], xs[1], xs[2]);
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            Y[3LL * nn + c] = yv[c];
            if (MODE == MODE_DOT && kFusedPcg && a.pout)
              reinterpret_cast<T*>(a.pout)[(long long)r * col + 3LL * nn + c] = xs[c];
          }
        }
      }
    }
    e_cur = e_next;
    (void)mk_next;
  }
  if (MODE == MODE_DOT || kChebDot || kWork) {  // partials of x.y (DOT), b.y (CHEB), r.r (RESID_WORK)
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
    if ((t & 63) == 0) red[t >> 6] = dot;
    __syncthreads();
    if (t == 0) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; ++w) s += red[w];
      a.partial[(long long)r * gridDim.x + blockIdx.x] = s;
    }
  }
}

struct Tiling {
  int TY, TZ, tiles_y, tiles_z, xchunk, chunks;
};

// compute units of the current device (the first query is kept: one device per process)
static int device_cus() {
  static int cus = 0;
  if (cus < 1) {
    int dev = 0, n = 0;
    cus = (hipGetDevice(&dev) == hipSuccess &&
           hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
              ? n
              : 256;
  }
  return cus;
}

static Tiling choose_tiling(const Geom& g, int k, int nt) {
  // z tile spans the whole node column when it fits (contiguous rows, full cache lines)
  Tiling best{};
  long long best_cost = -1;
  for (int TZ = 8; TZ <= std::min(nt / 4, 128); ++TZ) {
    const int TY = nt / TZ;
    if ((TY + 1) * (TZ + 1) > nt + 136 || TY < 4) continue;
    const int ty = (g.ny + TY - 2) / (TY - 1), tz = (g.nz + TZ - 2) / (TZ - 1);
    const long long cost = (long long)ty * tz * ((TY * TZ + 63) / 64) * 64;
    if (best_cost < 0 || cost < best_cost || (cost == best_cost && TZ > best.TZ)) {
      best_cost = cost;
      best = Tiling{TY, TZ, ty, tz, 0, 0};
    }
  }
  // x-chunks: each chunk recomputes one element layer and fills the plane pipeline, and a launch runs in rounds of
  // the workgroups the device holds at once (three per CU at the kernels' occupancy), so the chunk count with the
  // fewest rounds x (layers per chunk + 2) wins (ties: fewer chunks)
  const long long tiles = (long long)best.tiles_y * best.tiles_z * (k > 0 ? k : 1);
  const long long slots = 3LL * device_cus();
  long long best_t = -1;
  for (int c = 1; c <= std::max(1, g.nx / 16); ++c) {
    const int xc = (g.nx + c - 1) / c, cc = (g.nx + xc - 1) / xc;
    const long long t = ((tiles * cc + slots - 1) / slots) * (xc + 2);
    if (best_t < 0 || t < best_t) best_t = t, best.xchunk = xc, best.chunks = cc;
  }
  return best;
}

template <typename T>
int fine_grid_blocks(const Geom& g, int mode, int k) {
  const Tiling tl = choose_tiling(g, mode == MODE_DOT ? 1 : k, FineCfg<T>::NT);
  return tl.tiles_y * tl.tiles_z * tl.chunks;
}
template int fine_grid_blocks<float>(const Geom&, int, int);
template int fine_grid_blocks<double>(const Geom&, int, int);

int fine_blocks(const Geom& g) {
  Tiling a = choose_tiling(g, 1, FineCfg<float>::NT), b = choose_tiling(g, 1, FineCfg<double>::NT);
  return std::max(a.tiles_y * a.tiles_z * a.chunks, b.tiles_y * b.tiles_z * b.chunks);
}

template <typename T>
static KeSym<T> sym_blocks(const KeBlocks& kb) {
  KeSym<T> s;
  auto b = [&](int p, int i, int j) { return (T)kb.b[p][i][j]; };
  const double big = [&] {
    double m = 0;
    for (int p = 0; p < 8; ++p)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m = std::max(m, std::fabs(kb.b[p][i][j]));
    return m;
  }();
  // structural zeros of the axis-aligned brick (rigid-translation rows and the parity-forbidden couplings)
  const int zeros[][3] = {{1, 0, 0}, {1, 0, 1}, {1, 0, 2}, {1, 1, 0}, {1, 2, 0}, {2, 1, 0}, {2, 1, 1}, {2, 1, 2},
                          {2, 0, 1}, {2, 2, 1}, {4, 2, 0}, {4, 2, 1}, {4, 2, 2}, {4, 0, 2}, {4, 1, 2}, {3, 0, 2},
                          {3, 1, 2}, {3, 2, 0}, {3, 2, 1}, {5, 0, 1}, {5, 1, 0}, {5, 1, 2}, {5, 2, 1}, {6, 0, 1},
                          {6, 0, 2}, {6, 1, 0}, {6, 2, 0}};
  for (auto& z : zeros)
    if (std::fabs(kb.b[z[0]][z[1]][z[2]]) > 1e-10 * big)
      throw std::runtime_error("element stiffness does not have the block pattern of an axis-aligned brick");
  for (int p = 0; p < 8; ++p)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        if (std::fabs(kb.b[p][i][j] - kb.b[p][j][i]) > 1e-10 * big)
          throw std::runtime_error("element stiffness is not symmetric");
  const T v[33] = {b(0, 0, 0), b(0, 0, 1), b(0, 0, 2), b(0, 1, 1), b(0, 1, 2), b(0, 2, 2),
                   b(7, 0, 0), b(7, 0, 1), b(7, 0, 2), b(7, 1, 1), b(7, 1, 2), b(7, 2, 2),
                   b(1, 1, 1), b(1, 1, 2), b(1, 2, 2), b(2, 0, 0), b(2, 0, 2), b(2, 2, 2),
                   b(4, 0, 0), b(4, 0, 1), b(4, 1, 1), b(3, 0, 0), b(3, 0, 1), b(3, 1, 1), b(3, 2, 2),
                   b(5, 0, 0), b(5, 0, 2), b(5, 2, 2), b(5, 1, 1), b(6, 1, 1), b(6, 1, 2), b(6, 2, 2), b(6, 0, 0)};
  for (int i = 0; i < 33; ++i) s.k[i] = v[i];
  return s;
}

template <typename T>
void fine_apply(const Geom& g, const T* moduli, const uint8_t* mask, const KeBlocks& kbd, int mode,
                const ApplyArgs& a, hipStream_t s, int* nblocks_out) {
  if (g.nodes() * 3 > (long long)INT32_MAX) throw std::runtime_error("grid too large for 32-bit node indexing");
  constexpr int NT = FineCfg<T>::NT;
  Tiling tl = choose_tiling(g, mode == MODE_DOT ? 1 : a.k, NT);
  const KeSym<T> ks = sym_blocks<T>(kbd);
  dim3 grid(tl.tiles_y * tl.tiles_z * tl.chunks, a.k);
  if (nblocks_out) *nblocks_out = grid.x;
  // the premasked kernel addresses a column through 32-bit byte offsets
  const bool pre = a.premasked && g.nodes() * 3 * (long long)sizeof(T) < (1LL << 31);
#define DFMI_LAUNCH(M)                                                                                         \
  if (pre)                                                                                                     \
    hipLaunchKernelGGL((k_fine_apply<T, M, true>), grid, dim3(NT), 0, s, g, moduli, mask, ks, a, tl.TY, tl.TZ, \
                       tl.xchunk, tl.tiles_y, tl.tiles_z);                                                     \
  else                                                                                                         \
    hipLaunchKernelGGL((k_fine_apply<T, M, false>), grid, dim3(NT), 0, s, g, moduli, mask, ks, a, tl.TY,       \
                       tl.TZ, tl.xchunk, tl.tiles_y, tl.tiles_z)
  switch (mode) {
    case MODE_APPLY: DFMI_LAUNCH(MODE_APPLY); break;
    case MODE_RESID: DFMI_LAUNCH(MODE_RESID); break;
    case MODE_CHEB:
      if (sizeof(T) == sizeof(float) && a.partial) {
        DFMI_LAUNCH(kChebPartial);
      } else {
        DFMI_LAUNCH(MODE_CHEB);
      }
      break;
    case MODE_CHEB0: DFMI_LAUNCH(MODE_CHEB0); break;
    case MODE_CHEB1: DFMI_LAUNCH(MODE_CHEB1); break;
    case MODE_DOT: DFMI_LAUNCH(MODE_DOT); break;
    case MODE_MASS: DFMI_LAUNCH(MODE_MASS); break;
    case MODE_RESID_WORK:
      if constexpr (sizeof(T) == sizeof(double)) {
        if (!pre || !a.pout || !a.beta || !a.partial)
          throw std::runtime_error("fine_apply: RESID_WORK needs the premasked form, pout, beta and partials");
        hipLaunchKernelGGL((k_fine_apply<T, kResidWork, true>), grid, dim3(NT), 0, s, g, moduli, mask, ks, a, tl.TY,
                           tl.TZ, tl.xchunk, tl.tiles_y, tl.tiles_z);
      } else {
        throw std::runtime_error("fine_apply: RESID_WORK is an FP64 mode");
      }
      break;
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

// Sensitivity of the modal KS aggregate (topology_optimization.py:956-964): per element
// sum_r w_r (s_e u_re^T Ke u_re - lambda_r m_e u_re^T Me u_re) / target over k mode shapes, with
// Me = M8 (x) I3 diagonal in the corner-Hadamard basis (mdiag carries the 1/64 of H^T H = 8 I per axis).
struct ModalTerms {
  double weight[MODAL_MAX_MODES];
  double value[MODAL_MAX_MODES];
  double mdiag[8];
};

__global__ void k_modal_sensitivity(Geom pad, int ux, int uy, int uz, const double* __restrict__ u, int k,
                                    Blocks<double> kb, ModalTerms mt, const double* __restrict__ stiffness_slope,
                                    const double* __restrict__ mass_slope, double inv_target,
                                    double* __restrict__ out) {
  const long long ne = (long long)ux * uy * uz;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int el = (int)(e % uz);
  const int ej = (int)((e / uz) % uy);
  const int ei = (int)(e / ((long long)uz * uy));
  const double ss = stiffness_slope[e], ms = mass_slope[e];
  double acc = 0.0;
  for (int r = 0; r < k; ++r) {
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
    double strain = 0.0, kinetic = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double sq = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        strain = fma(v[q][c], w[q][c], strain);
        sq = fma(v[q][c], v[q][c], sq);
      }
      kinetic = fma(mt.mdiag[q], sq, kinetic);
    }
    acc += mt.weight[r] * (ss * strain - mt.value[r] * ms * kinetic);
  }
  out[e] = acc * inv_target;
}

void fine_modal_sensitivity(const Geom& pad, int ux, int uy, int uz, const double* u, int k, const KeBlocks& kbd,
                            const double* mdiag, const double* weight, const double* value,
                            const double* stiffness_slope, const double* mass_slope, double target, double* out,
                            hipStream_t s) {
  if (k < 1 || k > MODAL_MAX_MODES) throw std::invalid_argument("modal sensitivity takes 1 to 16 mode shapes");
  Blocks<double> kb;
  for (int p = 0; p < 8; ++p)
    for (int e = 0; e < 9; ++e) kb.b[p][e] = kbd.b[p][e / 3][e % 3];
  ModalTerms mt{};
  for (int r = 0; r < k; ++r) {
    mt.weight[r] = weight[r];
    mt.value[r] = value[r];
  }
  for (int q = 0; q < 8; ++q) mt.mdiag[q] = mdiag[q];
  const long long ne = (long long)ux * uy * uz;
  hipLaunchKernelGGL(k_modal_sensitivity, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, pad, ux, uy, uz, u, k,
                     kb, mt, stiffness_slope, mass_slope, 1.0 / target, out);
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
