// Fine-level hex8 operator on the matrix cores (gfx950 MFMA), FP32 in and out.
//
// For one right-hand side and one x-layer of elements, the element products Ke u_e of a tile of
// elements form one dense product Y = Ke [u_e1 u_e2 ...]: 24 x 24 by 24 x (elements).  It runs as
// 32 x 32 MFMA blocks of 32 elements, three bf16 products per block with FP32 accumulation,
//   Ke u ~ Ke_lo u_hi + Ke_hi u_lo + Ke_hi u_hi,
// where v_hi = bf16(v) and v_lo = bf16(v - v_hi).  That keeps 16 significant bits of both operands
// (relative error about 2^-16 per product instead of 2^-8 for a single bf16 product), enough for the
// fine-level smoother and residual of the FP32 V-cycle; the Krylov operator stays on the exact FP32
// kernel (fine.hip).
//
// Operand layouts (v_mfma_f32_32x32x16_bf16 for the x-planes' components 0 and 1, then
// v_mfma_f32_32x32x8bf16_1k for component 2):
//   * column n of a block = element (y, z) = (2 * wave + n / 16, 16 * zb + n % 16) of the tile;
//   * lane group g = lane / 32 supplies the K entries of node plane x = g of the element (components
//     0, 1 for the first product, component 2 for the second), the four (y, z) corners each;
//   * output row m = 8 c' + 4 x' + 2 y' + z' (rows 24..31 are zero), so the accumulator of lane group g
//     holds the element's contributions to its corners on node plane x = g.
// A workgroup owns a (TYE - 1)-row tile of node rows in y over the whole z extent (16 * ZB elements)
// and marches along x.  After each layer, the two z corners of every node are summed across lanes
// (DPP), both node planes of the layer go through LDS, and the owner of each node sums the two y
// corners and carries the next plane's part in registers.  Sums are in a fixed order, so results are
// deterministic.
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "dfmi_internal.h"

namespace dfmi {

typedef __attribute__((ext_vector_type(8))) __bf16 mf_bf16x8;
typedef __attribute__((ext_vector_type(4))) short mf_s16x4;
typedef __attribute__((ext_vector_type(16))) float mf_f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned mf_u32x4;
typedef __attribute__((ext_vector_type(2))) unsigned mf_u32x2;

// split v into bf16 hi (low half of the word) and bf16 lo = bf16(v - hi) (high half)
__device__ __forceinline__ unsigned mf_split(float v) {
  const __bf16 h = (__bf16)v;
  const __bf16 l = (__bf16)(v - (float)h);
  return (unsigned)__builtin_bit_cast(unsigned short, h) | ((unsigned)__builtin_bit_cast(unsigned short, l) << 16);
}
// bf16 pairs of two split words a, b: {hi(a), hi(b)} and {lo(a), lo(b)} (one v_perm_b32 each)
__device__ __forceinline__ unsigned mf_his(unsigned a, unsigned b) { return __builtin_amdgcn_perm(b, a, 0x05040100u); }
__device__ __forceinline__ unsigned mf_los(unsigned a, unsigned b) { return __builtin_amdgcn_perm(b, a, 0x07060302u); }
// value of lane (l - 1) mod 16 within each row of 16 lanes (DPP row_ror:1)
__device__ __forceinline__ float mf_prev_lane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x121, 0xF, 0xF, true));
}

template <int MODE, int TYE, int ZB>
__global__ __launch_bounds__(TYE * 32) __attribute__((amdgpu_waves_per_eu(ZB <= 3 ? 4 : 2))) void k_fine_mfma(Geom g, const float* __restrict__ moduli,
                                                        const uint8_t* __restrict__ mask,
                                                        const unsigned* __restrict__ atab, ApplyArgs a, int xchunk,
                                                        int tiles_y) {
  constexpr int NT = TYE * 32;          // TYE / 2 waves, one element-row pair each
  constexpr int SZ = ZB * 16 + 1;       // node slots along z
  constexpr int NS = (TYE + 1) * SZ;    // node slots of a staged plane (one halo row at the low side)
  constexpr int KST = (NS + NT - 1) / NT;
  constexpr int NOWN = (TYE - 1) * SZ;  // owned node slots of a plane
  constexpr int KEP = (NOWN + NT - 1) / NT;
  constexpr bool NEEDX = MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_CHEB1;
  // LDS strides padded so that the four 16-lane groups of a wave (two element rows, two node planes) fall
  // on different banks: plane buffers 32 words apart modulo 64, accumulator groups likewise
  constexpr int NSP = NS + ((32 - NS % 64) + 64) % 64;  // Pk stride per component, = 32 (mod 64)
  constexpr int NSA = NS + ((16 - NS % 32) + 32) % 32;  // ACC stride per (y corner, component), = 16 (mod 32)
  __shared__ unsigned Pk[2][3][NSP];   // node planes i, i+1 of the operand, split into bf16 hi / lo
  __shared__ float ACC[2][2][3][NSA];  // [plane i / i+1][y corner][component][node slot], z already summed

  const int r = blockIdx.y;
  if (a.active && !a.active[r]) return;
  int tile = (int)xcd_block(blockIdx.x, gridDim.x);
  const int ty = tile % tiles_y;
  const int cx = tile / tiles_y;
  const int y0 = ty * (TYE - 1);
  const int i0 = cx * xchunk;
  const int i1 = min(i0 + xchunk, g.nx);
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, n = lane & 31, grp = lane >> 5;
  const int ye = 2 * wave + (n >> 4), zl = n & 15;
  const long long col = (long long)g.nodes() * 3;
  const float* X = reinterpret_cast<const float*>(a.x) + (long long)r * col;
  const float* B = reinterpret_cast<const float*>(a.b) + (long long)r * col;
  const float* XP = reinterpret_cast<const float*>(a.xp) + (long long)r * col;
  const float* DI = reinterpret_cast<const float*>(a.dinv);
  float* Y = reinterpret_cast<float*>(a.y) + (long long)r * col;
  const int plane_nodes = g.ny * g.nz;

  // A operands (Ke split into bf16 hi / lo), fixed for the kernel
  mf_u32x4 ahb, alb;
  mf_u32x2 ahs, als;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    ahb[j] = atab[lane * 12 + j];
    alb[j] = atab[lane * 12 + 4 + j];
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    ahs[j] = atab[lane * 12 + 8 + j];
    als[j] = atab[lane * 12 + 10 + j];
  }
  const mf_bf16x8 Ahb = __builtin_bit_cast(mf_bf16x8, ahb), Alb = __builtin_bit_cast(mf_bf16x8, alb);
  const mf_s16x4 Ahs = __builtin_bit_cast(mf_s16x4, ahs), Als = __builtin_bit_cast(mf_s16x4, als);

  // staging of node plane ip (slots t + k NT); so[k]: node offset within a plane, or -1 outside the grid
  int so[KST];
#pragma unroll
  for (int k = 0; k < KST; ++k) {
    const int v = t + k * NT;
    const int yy = v / SZ, zz = v - (v / SZ) * SZ;
    const int gy = y0 - 1 + yy;
    so[k] = (v < NS && gy >= 0 && gy < g.ny && zz < g.nz) ? gy * g.nz + zz : -1;
  }
  unsigned sv[KST][3];
  auto fetch = [&](int ip) {
    const bool plane_in = ip >= 0 && ip < g.nx;
#pragma unroll
    for (int k = 0; k < KST; ++k) {
      float v0 = 0.f, v1 = 0.f, v2 = 0.f;
      {
        if (plane_in && so[k] >= 0) {
          const int nn = ip * plane_nodes + so[k];
          const uint8_t m = mask[nn];
          if (MODE == MODE_CHEB0) {
            const float c0 = (float)a.c0;
            v0 = c0 * DI[3 * nn] * B[3 * nn];
            v1 = c0 * DI[3 * nn + 1] * B[3 * nn + 1];
            v2 = c0 * DI[3 * nn + 2] * B[3 * nn + 2];
          } else {
            v0 = X[3 * nn];
            v1 = X[3 * nn + 1];
            v2 = X[3 * nn + 2];
          }
          v0 = (m & 1) ? v0 : 0.f;
          v1 = (m & 2) ? v1 : 0.f;
          v2 = (m & 4) ? v2 : 0.f;
        }
      }
      sv[k][0] = mf_split(v0);
      sv[k][1] = mf_split(v1);
      sv[k][2] = mf_split(v2);
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int k = 0; k < KST; ++k) {
      const int v = t + k * NT;
      if (v >= NS) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) Pk[buf][c][v] = sv[k][c];
    }
  };
  // moduli of this lane's elements in layer i
  float en[ZB];
  auto modulus = [&](int i) {
    const int gy = y0 - 1 + ye;
#pragma unroll
    for (int zb = 0; zb < ZB; ++zb) {
      const int ze = 16 * zb + zl;
      en[zb] = (i >= 0 && i < g.ex && gy >= 0 && gy < g.ey && ze < g.ez)
                   ? moduli[(i * g.ey + gy) * g.ez + ze]
                   : 0.f;
    }
  };

  // owned nodes of this thread (same slots on every plane) and their carried x-high contributions
  int own[KEP];
  float ycarry[KEP][3];
#pragma unroll
  for (int k = 0; k < KEP; ++k) {
    const int v = t + k * NT;
    own[k] = -1;
    ycarry[k][0] = ycarry[k][1] = ycarry[k][2] = 0.f;
    if (v < NOWN) {
      const int yy = 1 + v / SZ, zz = v - (v / SZ) * SZ;
      const int gy = y0 - 1 + yy;
      if (gy < g.ny && zz < g.nz) own[k] = gy * g.nz + zz;
    }
  }

  fetch(i0 - 1);
  stage((i0 - 1) & 1);
  fetch(i0);
  modulus(i0 - 1);
  for (int i = i0 - 1; i < i1; ++i) {
    stage((i + 1) & 1);
    if (i + 2 <= i1) fetch(i + 2);
    float ecur[ZB];
#pragma unroll
    for (int zb = 0; zb < ZB; ++zb) ecur[zb] = en[zb];
    if (i + 1 < i1) modulus(i + 1);
    const bool out = i >= i0;
    // epilogue operands of the owned nodes of plane i, requested early
    float bv[KEP][3], xpv[KEP][3], dv[KEP][3], xs[KEP][3];
    uint8_t mk[KEP];
#pragma unroll
    for (int k = 0; k < KEP; ++k) {
      mk[k] = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) bv[k][c] = xpv[k][c] = dv[k][c] = xs[k][c] = 0.f;
      if (!out || own[k] < 0) continue;
      const int nn = i * plane_nodes + own[k];
      mk[k] = mask[nn];
      if (MODE == MODE_RESID || NEEDX) {
#pragma unroll
        for (int c = 0; c < 3; ++c) bv[k][c] = B[3 * nn + c];
      }
      if (NEEDX) {
#pragma unroll
        for (int c = 0; c < 3; ++c) dv[k][c] = DI[3 * nn + c];
        if (MODE == MODE_CHEB0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) xs[k][c] = ((mk[k] >> c) & 1) ? (float)a.c0 * dv[k][c] * bv[k][c] : 0.f;
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) xs[k][c] = X[3 * nn + c];
        }
        if (MODE == MODE_CHEB && a.xp) {
#pragma unroll
          for (int c = 0; c < 3; ++c) xpv[k][c] = XP[3 * nn + c];
        }
      }
    }
    __syncthreads();  // planes i and i+1 staged; the previous plane's epilogue is done with ACC
    const int pb = (i + grp) & 1;
    mf_f32x16 acc[ZB];
#pragma unroll
    for (int zb = 0; zb < ZB; ++zb) {
      const int s00 = ye * SZ + 16 * zb + zl;
      unsigned d[3][4];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        d[c][0] = Pk[pb][c][s00];
        d[c][1] = Pk[pb][c][s00 + 1];
        d[c][2] = Pk[pb][c][s00 + SZ];
        d[c][3] = Pk[pb][c][s00 + SZ + 1];
      }
      mf_u32x4 bhb, blb;
      mf_u32x2 bhs, bls;
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int y2 = 0; y2 < 2; ++y2) {
          bhb[2 * c + y2] = mf_his(d[c][2 * y2], d[c][2 * y2 + 1]);
          blb[2 * c + y2] = mf_los(d[c][2 * y2], d[c][2 * y2 + 1]);
        }
#pragma unroll
      for (int y2 = 0; y2 < 2; ++y2) {
        bhs[y2] = mf_his(d[2][2 * y2], d[2][2 * y2 + 1]);
        bls[y2] = mf_los(d[2][2 * y2], d[2][2 * y2 + 1]);
      }
      const mf_bf16x8 Bhb = __builtin_bit_cast(mf_bf16x8, bhb), Blb = __builtin_bit_cast(mf_bf16x8, blb);
      const mf_s16x4 Bhs = __builtin_bit_cast(mf_s16x4, bhs), Bls = __builtin_bit_cast(mf_s16x4, bls);
      mf_f32x16 c0v = {};
      c0v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Alb, Bhb, c0v, 0, 0, 0);
      c0v = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(Als, Bhs, c0v, 0, 0, 0);
      c0v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahb, Blb, c0v, 0, 0, 0);
      c0v = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(Ahs, Bls, c0v, 0, 0, 0);
      c0v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahb, Bhb, c0v, 0, 0, 0);
      acc[zb] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(Ahs, Bhs, c0v, 0, 0, 0);
    }
    // scale by the element moduli, sum the two z corners of each node within the wave (the z-high corner
    // of element z is the node of element z + 1: lane zl + 1, or lane 0 of the next block for zl = 15),
    // and store per (plane, y corner, component); lane group 1 holds the contributions to plane i+1
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int y2 = 0; y2 < 2; ++y2) {
        float* dst = &ACC[grp][y2][c][(ye + y2) * SZ];
        float rot[ZB];
#pragma unroll
        for (int zb = 0; zb < ZB; ++zb) rot[zb] = mf_prev_lane(acc[zb][4 * c + 2 * y2 + 1] * ecur[zb]);
#pragma unroll
        for (int zb = 0; zb < ZB; ++zb) {
          const float lo = acc[zb][4 * c + 2 * y2] * ecur[zb];
          const float prev = zl == 0 ? (zb > 0 ? rot[zb - 1] : 0.f) : rot[zb];
          dst[16 * zb + zl] = lo + prev;
        }
        if (zl == 0) dst[16 * ZB] = rot[ZB - 1];
      }
    __syncthreads();  // ACC holds every contribution to planes i and i+1 of layer i
#pragma unroll
    for (int k = 0; k < KEP; ++k) {
      if (own[k] < 0) continue;
      const int s = t + k * NT + SZ;
      float sum[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sum[c] = ycarry[k][c] + (ACC[0][0][c][s] + ACC[0][1][c][s]);
        ycarry[k][c] = ACC[1][0][c][s] + ACC[1][1][c][s];
      }
      if (!out) continue;
      const int nn = i * plane_nodes + own[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const bool on = (mk[k] >> c) & 1;
        float yv;
        if (MODE == MODE_APPLY) {
          yv = sum[c];
        } else if (MODE == MODE_RESID) {
          yv = bv[k][c] - sum[c];
        } else {
          const float c1 = (float)a.c1, c2 = (float)a.c2;
          const float prev = MODE == MODE_CHEB1 ? (float)a.c0 * dv[k][c] * bv[k][c] : xpv[k][c];
          yv = xs[k][c] + c1 * (xs[k][c] - prev) + c2 * dv[k][c] * (bv[k][c] - sum[c]);
        }
        Y[3 * nn + c] = on ? yv : 0.f;
      }
    }
  }
}

// ---- host side -------------------------------------------------------------------------------

static uint16_t bf16_bits(float v) {  // round to nearest even
  uint32_t u;
  std::memcpy(&u, &v, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float v;
  std::memcpy(&v, &u, 4);
  return v;
}

// Per-lane A operands of the two MFMA shapes: 12 words per lane (hi 32x16, lo 32x16, hi 32x8, lo 32x8).
// ke_bits: 24 x 24 element stiffness, DOF 3 q + c with corner bits q = x + 2 y + 4 z.
void fine_mfma_table(const double* ke_bits, std::vector<unsigned>& out) {
  out.assign(64 * 12, 0u);
  auto ke = [&](int m, int x, int c, int y, int z) -> double {
    const int cp = m / 8, xp = (m / 4) & 1, yp = (m / 2) & 1, zp = m & 1;
    if (cp >= 3) return 0.0;
    const int qp = xp + 2 * yp + 4 * zp, q = x + 2 * y + 4 * z;
    return ke_bits[(3 * qp + cp) * 24 + 3 * q + c];
  };
  for (int lane = 0; lane < 64; ++lane) {
    const int m = lane % 32, grp = lane / 32;
    uint16_t hb[8], lb[8], hs[4], ls[4];
    for (int j = 0; j < 8; ++j) {
      const double v = ke(m, grp, j / 4, (j % 4) / 2, j % 2);
      hb[j] = bf16_bits((float)v);
      lb[j] = bf16_bits((float)(v - (double)bf16_value(hb[j])));
    }
    for (int j = 0; j < 4; ++j) {
      const double v = ke(m, grp, 2, j / 2, j % 2);
      hs[j] = bf16_bits((float)v);
      ls[j] = bf16_bits((float)(v - (double)bf16_value(hs[j])));
    }
    unsigned* w = out.data() + lane * 12;
    for (int j = 0; j < 4; ++j) {
      w[j] = hb[2 * j] | ((unsigned)hb[2 * j + 1] << 16);
      w[4 + j] = lb[2 * j] | ((unsigned)lb[2 * j + 1] << 16);
    }
    for (int j = 0; j < 2; ++j) {
      w[8 + j] = hs[2 * j] | ((unsigned)hs[2 * j + 1] << 16);
      w[10 + j] = ls[2 * j] | ((unsigned)ls[2 * j + 1] << 16);
    }
  }
}

bool fine_mfma_supported(const Geom& g, int mode) {
  const bool mode_ok = mode == MODE_APPLY || mode == MODE_RESID || mode == MODE_CHEB || mode == MODE_CHEB0 ||
                       mode == MODE_CHEB1;
  return mode_ok && g.ez >= 1 && g.ez <= 16 * FINE_MFMA_ZMAX && g.nodes() * 3 <= (long long)INT32_MAX;
}

template <int MODE, int TYE, int ZB>
static void launch_tye_zb(const Geom& g, const float* moduli, const uint8_t* mask, const unsigned* atab,
                          const ApplyArgs& a, const FineMfmaTile& tl, hipStream_t s) {
  const int tiles_y = (g.ny + TYE - 2) / (TYE - 1);  // TYE - 1 owned node rows per tile
  const int chunks = (g.nx + tl.xchunk - 1) / tl.xchunk;
  dim3 grid(tiles_y * chunks, a.k);
  hipLaunchKernelGGL((k_fine_mfma<MODE, TYE, ZB>), grid, dim3(TYE * 32), 0, s, g, moduli, mask, atab, a, tl.xchunk,
                     tiles_y);
}

template <int MODE, int TYE>
static void launch_tye(const Geom& g, const float* moduli, const uint8_t* mask, const unsigned* atab,
                       const ApplyArgs& a, const FineMfmaTile& tl, hipStream_t s) {
  switch ((g.ez + 15) / 16) {
    case 1: launch_tye_zb<MODE, TYE, 1>(g, moduli, mask, atab, a, tl, s); break;
    case 2: launch_tye_zb<MODE, TYE, 2>(g, moduli, mask, atab, a, tl, s); break;
    case 3: launch_tye_zb<MODE, TYE, 3>(g, moduli, mask, atab, a, tl, s); break;
    case 4: launch_tye_zb<MODE, TYE, 4>(g, moduli, mask, atab, a, tl, s); break;
    default: throw std::runtime_error("fine_apply_mfma: z extent beyond 64 elements");
  }
}

template <int MODE>
static void launch_mode(const Geom& g, const float* moduli, const uint8_t* mask, const unsigned* atab,
                        const ApplyArgs& a, const FineMfmaTile& tl, hipStream_t s) {
  if (tl.tye == 16)
    launch_tye<MODE, 16>(g, moduli, mask, atab, a, tl, s);
  else if (tl.tye == 8)
    launch_tye<MODE, 8>(g, moduli, mask, atab, a, tl, s);
  else
    throw std::runtime_error("fine_apply_mfma: tile rows must be 8 or 16");
}

void fine_apply_mfma(const Geom& g, const float* moduli, const uint8_t* mask, const unsigned* atab, int mode,
                     const ApplyArgs& a, const FineMfmaTile& tl, hipStream_t s) {
  if (!fine_mfma_supported(g, mode)) throw std::runtime_error("fine_apply_mfma: unsupported grid or mode");
  if (tl.xchunk < 1) throw std::runtime_error("fine_apply_mfma: bad x chunk");
  switch (mode) {
    case MODE_APPLY: launch_mode<MODE_APPLY>(g, moduli, mask, atab, a, tl, s); break;
    case MODE_RESID: launch_mode<MODE_RESID>(g, moduli, mask, atab, a, tl, s); break;
    case MODE_CHEB: launch_mode<MODE_CHEB>(g, moduli, mask, atab, a, tl, s); break;
    case MODE_CHEB0: launch_mode<MODE_CHEB0>(g, moduli, mask, atab, a, tl, s); break;
    case MODE_CHEB1: launch_mode<MODE_CHEB1>(g, moduli, mask, atab, a, tl, s); break;
    default: throw std::runtime_error("fine_apply_mfma: bad mode");
  }
  DFMI_HIP(hipGetLastError());
}

// Candidate tiles for the autotuner: tile rows (8 or 16) and x-chunk lengths that give a few
// thousand workgroups at this column count.
std::vector<FineMfmaTile> fine_mfma_candidates(const Geom& g, int k) {
  std::vector<FineMfmaTile> out;
  for (int tye : {16, 8}) {
    const int tiles_y = (g.ny + tye - 2) / (tye - 1);
    for (int target : {1024, 2048, 4096, 8192}) {
      long long tiles = (long long)tiles_y * std::max(1, k);
      int chunks = (int)std::max(1LL, (target + tiles - 1) / tiles);
      chunks = std::min(chunks, std::max(1, g.nx / 8));
      const int xchunk = (g.nx + chunks - 1) / chunks;
      bool dup = false;
      for (auto& c : out) dup = dup || (c.tye == tye && c.xchunk == xchunk);
      if (!dup) out.push_back(FineMfmaTile{tye, xchunk});
    }
  }
  return out;
}

}  // namespace dfmi
