// Internal declarations shared by the libdfmi translation units.
// Grid conventions (same numbering as deep_frame/topology_optimization.py:301-319):
//   node (i, j, l) of an (X+1, Y+1, Z+1) node grid has index (i*(Y+1) + j)*(Z+1) + l,
//   element (i, j, l) of an (X, Y, Z) element grid has index (i*Y + j)*Z + l,
//   a multi-vector is stored RHS-major: v[r][node][comp], comp fastest.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

namespace dfmi {

void set_error(const std::string& msg);
const char* get_error();

#define DFMI_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_));  \
    }                                                                               \
  } while (0)

// Node/element geometry of one level.
struct Geom {
  int ex, ey, ez;   // elements per axis
  int nx, ny, nz;   // nodes per axis (= elements + 1)
  __host__ __device__ long long nodes() const { return (long long)nx * ny * nz; }
  __host__ __device__ long long elems() const { return (long long)ex * ey * ez; }
};

inline Geom make_geom(int ex, int ey, int ez) { return Geom{ex, ey, ez, ex + 1, ey + 1, ez + 1}; }

// Multi-vector layout: entry (column r, node n, component c) lives at n * ns + r * rs + c.
// The fine level is column-major (one contiguous column per right-hand side, as the flat DOF
// vectors of the interface); coarse levels interleave the right-hand sides per node so that a
// stencil gather of one neighbour is a single contiguous run for all columns of a thread.
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  Remap the hardware
// block index so that every XCD works on one contiguous range of logical blocks (neighbouring
// tiles share halo data in the same L2).
__device__ inline unsigned xcd_block(unsigned b, unsigned nblocks) {
  constexpr unsigned XCD = 8;
  const unsigned per = (nblocks + XCD - 1) / XCD;
  const unsigned x = b % XCD, i = b / XCD;
  const unsigned full = nblocks % XCD == 0 ? XCD : nblocks % XCD;  // XCDs that get `per` blocks
  // XCDs [0, full) own `per` blocks each, the others own per - 1
  return x < full ? x * per + i : full * per + (x - full) * (per - 1) + i;
}

// Block-wide reductions for 256-thread workgroups (four 64-wide wavefronts).
__device__ inline double block_sum(double v) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__device__ inline double block_max(double v) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

struct VLay {
  long long ns, rs;
};
inline VLay column_major(const Geom& g) { return VLay{3, g.nodes() * 3}; }
inline VLay interleaved(int k) { return VLay{3LL * k, 3}; }

// Hex8 stiffness in the corner-Hadamard basis: 8 parity classes of 3x3 blocks,
// pre-scaled by 1/64 (see hex8_blocks() in the host code).
struct KeBlocks {
  double b[8][3][3];
};

// Apply epilogues shared by the fine (matrix-free) and coarse (stencil) kernels.
enum ApplyMode : int {
  MODE_APPLY = 0,   // y = M A M x
  MODE_RESID = 1,   // y = M (b - A M x)
  MODE_CHEB = 2,    // y = x + c1 (x - xp) + c2 Dinv (b - A x)   (y may alias xp)
  MODE_DOT = 3,     // y = M A M x and partial sums of x.y per workgroup
  MODE_ENERGY = 4,  // fine only: per-element u_e^T Ke u_e (no modulus)
  MODE_CHEB0 = 5,   // fine only: first Chebyshev step from zero, the operand x1 = c0 Dinv b is formed
                    // while staging (a.x unused): y = x1 + c1 x1 + c2 Dinv (b - A x1)
  MODE_CHEB1 = 6,   // fine only: Chebyshev step whose previous iterate is x1 = c0 Dinv b (a.xp unused)
  MODE_MASS = 7,    // fine only: y = M (sum_e m_e Me + diag(lumped)) M x with the element field m_e in place of
                    // the moduli, Me from a.mdiag (Hadamard-diagonal form) and lumped in a.dinv (may be null)
  MODE_RESID_WORK = 8,  // fine FP64 only: r = M (b - A M x) written in FP32 as r * beta[c] to a.pout (a.y unused),
                        // with per-workgroup partial sums of |r|^2 (FP64) in a.partial
};

// Three consecutive entries of a vector through a raw buffer load: byte offsets at or beyond the buffer's
// size read zeros (no clamp or select needed for neighbours or slots outside the array), and the address is
// one 32-bit offset from a scalar base.  Buffers must stay below 2 GiB.
__device__ __forceinline__ void load3(__amdgpu_buffer_rsrc_t r, unsigned byte, float& a, float& b, float& c) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, byte, 0, 0);
  a = __uint_as_float(v[0]);
  b = __uint_as_float(v[1]);
  c = __uint_as_float(v[2]);
}
__device__ __forceinline__ void load3(__amdgpu_buffer_rsrc_t r, unsigned byte, double& a, double& b, double& c) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, byte, 0, 0);
  const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, byte + 16, 0, 0);
  a = __hiloint2double((int)v[1], (int)v[0]);
  b = __hiloint2double((int)v[3], (int)v[2]);
  c = __hiloint2double((int)w[1], (int)w[0]);
}
// buffer descriptor word 3 for raw 32-bit loads on gfx9 (data format 32)
constexpr int kBufferWord3 = 0x00020000;
// a byte offset that lies beyond every buffer built by buffer() (sizes below 2 GiB)
constexpr unsigned kOutside = 0x80000000u;

// Three consecutive entries written through one (FP32) or two (FP64) buffer stores.
__device__ __forceinline__ void store3(__amdgpu_buffer_rsrc_t r, unsigned byte, float a, float b, float c) {
  __attribute__((ext_vector_type(3))) unsigned v = {__float_as_uint(a), __float_as_uint(b), __float_as_uint(c)};
  __builtin_amdgcn_raw_buffer_store_b96(v, r, byte, 0, 0);
}
__device__ __forceinline__ void store3(__amdgpu_buffer_rsrc_t r, unsigned byte, double a, double b, double c) {
  __attribute__((ext_vector_type(4))) unsigned v = {(unsigned)__double2loint(a), (unsigned)__double2hiint(a),
                                                    (unsigned)__double2loint(b), (unsigned)__double2hiint(b)};
  __attribute__((ext_vector_type(2))) unsigned w = {(unsigned)__double2loint(c), (unsigned)__double2hiint(c)};
  __builtin_amdgcn_raw_buffer_store_b128(v, r, byte, 0, 0);
  __builtin_amdgcn_raw_buffer_store_b64(w, r, byte + 16, 0, 0);
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer(const void* p, long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kBufferWord3);
}

struct ApplyArgs {
  const void* x;      // operand, [k][nodes][3]
  void* y;            // output
  const void* b;      // right-hand side for RESID / CHEB
  const void* xp;     // previous iterate for CHEB
  const void* dinv;   // [nodes][3] inverse diagonal (masked), CHEB only
  double c1, c2;      // Chebyshev coefficients, CHEB only
  double c0;          // first Chebyshev iterate x1 = c0 Dinv b, CHEB0 / CHEB1 only
  double* partial;    // [k][nblocks] partial dot products: x.y for DOT; b.y for FP32 fine CHEB when not null
  const double* beta; // FP32 fine DOT with pout: per-column beta of the fused direction (nullptr: p = b);
                      // FP64 fine RESID_WORK: per-column factor of the FP32 residual
  void* pout;         // FP32 fine DOT: when not null the operand is p = b + beta x, written here for owned nodes;
                      // FP64 fine RESID_WORK: the FP32 residual (same column stride as x)
  const int* active;  // per-column flag (device), nullptr = all active
  int k;
  double mdiag[8];    // MASS only: unit-density element mass in the corner-Hadamard basis, divided by 64
  int premasked;      // fine vector kernel: the operand (and for CHEB0 / DOT-with-pout the vectors the operand is
                      // formed from) is already zero on every DOF the mask cuts, so staging skips the mask
};

// ---- fine level (matrix-free hex8) -------------------------------------------------
template <typename T>
void fine_apply(const Geom& g, const T* moduli, const uint8_t* mask, const KeBlocks& kb, int mode,
                const ApplyArgs& a, hipStream_t s, int* nblocks_out);
int fine_blocks(const Geom& g);
// workgroups per column of a fine_apply launch in `mode` for k columns (the stride of its partials)
template <typename T>
int fine_grid_blocks(const Geom& g, int mode, int k);
// Fine level on the matrix cores (FP32 vectors, split-bf16 MFMA products): modes APPLY, RESID, CHEB,
// CHEB0, CHEB1 on grids with at most 16 * FINE_MFMA_ZMAX elements along z.  atab: per-lane A operands
// from fine_mfma_table (64 * 12 words, device).
constexpr int FINE_MFMA_ZMAX = 4;
struct FineMfmaTile {
  int tye;     // element rows per workgroup tile (8 or 16); TYE - 1 node rows are owned
  int xchunk;  // node planes per workgroup along x
};
void fine_mfma_table(const double* ke_bits, std::vector<unsigned>& out);
bool fine_mfma_supported(const Geom& g, int mode);
void fine_apply_mfma(const Geom& g, const float* moduli, const uint8_t* mask, const unsigned* atab, int mode,
                     const ApplyArgs& a, const FineMfmaTile& tile, hipStream_t s);
std::vector<FineMfmaTile> fine_mfma_candidates(const Geom& g, int k);
void fine_element_energy(const Geom& pad, int ux, int uy, int uz, const double* u, const KeBlocks& kb,
                         double* energy, int k, hipStream_t s);
constexpr int MODAL_MAX_MODES = 16;
// per element sum_r weight_r (stiffness_slope_e u_re^T Ke u_re - value_r mass_slope_e u_re^T Me u_re) / target
void fine_modal_sensitivity(const Geom& pad, int ux, int uy, int uz, const double* u, int k, const KeBlocks& kb,
                            const double* mdiag, const double* weight, const double* value,
                            const double* stiffness_slope, const double* mass_slope, double target, double* out,
                            hipStream_t s);
void fine_diagonal(const Geom& g, const double* moduli, const uint8_t* mask, const double* ke_diag, double* diag,
                   hipStream_t s);
void fine_stencil(const Geom& g, const double* moduli, const uint8_t* mask, const double* ke, double* stencil,
                  hipStream_t s);
void active_mask(const Geom& g, const double* moduli, const uint8_t* fixed, uint8_t* mask, hipStream_t s);
// same with the fixed flags given per DOF of the unpadded (ux+1, uy+1, uz+1) node grid
void active_mask_flat(const Geom& pad, int ux, int uy, int uz, const double* moduli, const uint8_t* fixed,
                      uint8_t* mask, hipStream_t s);

// ---- coarse levels (27-point block stencils, SoA [27*9][nodes]) -----------------------
template <typename T>
void stencil_apply(const Geom& g, const T* st, const uint8_t* mask, int mode, const ApplyArgs& a, VLay lay,
                   hipStream_t s, int* nblocks_out);
int stencil_blocks(const Geom& g, int k);
void stencil_rap(const Geom& fine, const double* fst, const Geom& coarse, double* cst, hipStream_t s);
void stencil_diagonal(const Geom& g, const double* st, double* diag, hipStream_t s);
void stencil_mask(const Geom& g, double* st, const uint8_t* mask, hipStream_t s);
void dead_mask(const Geom& g, const double* diag, double threshold, uint8_t* mask, hipStream_t s);
template <typename T>
void restrict_full(const Geom& fine, const T* f, VLay flay, const Geom& coarse, const uint8_t* cmask, T* c,
                   VLay clay, int k, const int* active, hipStream_t s);
template <typename T>
void prolong_add(const Geom& coarse, const T* c, VLay clay, const Geom& fine, const uint8_t* fmask, T* f,
                 VLay flay, int k, const int* active, hipStream_t s);
// interleaved [node][r][3] <-> column-major [r][node][3] (to_columns: interleaved -> column-major)
template <typename T>
void relayout(const T* in, T* out, long long nodes, int k, bool to_columns, const int* active, hipStream_t s);
void coarsest_assemble(const Geom& g, const double* st, const int* live_index, int n, double* dense, hipStream_t s);
void dense_sweep_inverse(double* a, int n, hipStream_t s);
// work_in, work_out: k * n doubles; work_split: coarsest_splits() * k * n doubles
template <typename T>
void coarsest_solve(const Geom& g, const double* ainv, const int* live_dofs, int n, const T* rhs, T* out, VLay lay,
                    double* work_in, double* work_out, double* work_split, int k, const int* active, hipStream_t s);
int coarsest_splits();

// ---- vectors ----------------------------------------------------------------------------
void reduce_partials(const double* partial, int nblocks, int k, double* out, hipStream_t s);
template <typename T>
void column_dot(const T* a, const T* b, long long n, int k, double* partial, int nblocks, const int* active,
                hipStream_t s);
template <typename T>
void column_absmax(const T* a, long long n, int k, double* partial, int nblocks, const int* active, hipStream_t s);
void reduce_max(const double* partial, int nblocks, int k, double* out, hipStream_t s);
template <typename T>
void cheb_start(const T* b, const T* dinv, double c, T* x, long long nodes, int k, const int* active, VLay lay,
                hipStream_t s);
void power_init(const uint8_t* mask, double* v, long long nodes, unsigned seed, hipStream_t s);
void inv_sqrt(const double* in, double* out, hipStream_t s);  // out[0] = 1 / sqrt(in[0])
// v = dinv * y * inv_norm[0] (dinv null: 1; inv_norm null: 1)
void dinv_scale(const double* y, const double* dinv, const double* inv_norm, double* v, long long nodes,
                hipStream_t s);
void pcg_beta(const double* rzn, const double* rzc, double* rz, double* beta, int k, const int* active,
              hipStream_t s);
template <typename T>
void copy_columns(const T* src, T* dst, long long n, int k, const int* select, hipStream_t s);
template <typename T>
void fill_zero(T* x, long long n, hipStream_t s);
// partial_rr: [k][nblocks] sums of r.r, then (z != nullptr) [k][nblocks] sums of r.z
template <typename T>
void pcg_update(T* x, T* r, const T* p, const T* ap, const double* alpha, long long n, int k, const int* active,
                double* partial_rr, int nblocks, hipStream_t s, const T* z = nullptr);
template <typename T>
void pcg_direction(T* p, const T* z, const double* beta, long long n, int k, const int* active, hipStream_t s);
// x += scale d on the active columns; with zero_d the FP32 increment d of those columns is then set to zero
void add_scaled(double* x, float* d, const double* scale, long long n, int k, const int* active, hipStream_t s,
                bool zero_d = false);
// Krylov coarse solve on interleaved level vectors [node][r][3]: dots of np (2 or 4) pairs a[i].b[i] per
// column into partial[(i * k + r) * nblocks + block]; scalar steps on the rows of ks (9 k doubles); y -= f x;
// x = a x + b p
template <typename T>
struct LevelPairs {
  const T* a[4];
  const T* b[4];
};
int level_dot_blocks(long long nodes);
template <typename T>
void level_dots(const LevelPairs<T>& p, int np, long long nodes, int k, const int* active, double* partial,
                int nblocks, hipStream_t s);
void krylov_scalars(double* ks, int k, const int* active, int step, hipStream_t s);
template <typename T>
void level_axpy(T* y, const T* x, const double* f, long long nodes, int k, const int* active, hipStream_t s);
template <typename T>
void level_combine(T* x, const T* p, const double* a, const double* b, long long nodes, int k, const int* active,
                   hipStream_t s);
// v_c *= f[c] on the selected columns; a zero factor, or f == nullptr, stores zeros without reading v
template <typename T>
void scale_columns(T* v, long long n, int k, const double* f, const int* select, hipStream_t s);
// dst[c] = src[c] (square_scale false) or dst[c] *= src[c]^2 (true) for selected columns
void select_scalars(const double* src, double* dst, const int* select, int k, bool square_scale, hipStream_t s);
void pcg_scalars_alpha(const double* rz, const double* pap, double* alpha, int k, const int* active, hipStream_t s);
template <typename T>
void to_work(const double* r, const double* scale, T* w, long long n, int k, const int* active, hipStream_t s);
template <typename T>
void from_work(const T* w, const double* scale, double* z, long long n, int k, const int* active, hipStream_t s);
// out = scale * A B^T for row blocks A (ka x n), B (kb x n), 1 <= ka, kb <= 64 (partial sums per workgroup
// in `work`, work_blocks * ka * kb doubles, summed in a fixed order)
void gram_device(const double* a, const double* b, long long n, int ka, int kb, double* out, double* work,
                 int work_blocks, double scale, hipStream_t s);
// passes of the block PCG recurrence over FP32 column-major multi-vectors (block.hip); coefficient matrices
// of block_combine / block_update are FP32 with row stride 48
int block_gram_blocks(long long n);
int block_combine_blocks(long long n);
void block_gram(const float* a, int ka, const float* b1, int kb1, const float* b2, int kb2, long long n, double* out,
                double* work, hipStream_t s);
void block_combine(const float* u, const float* x, float* y, const float* m, int kin, int kout, long long n,
                   hipStream_t s);
void block_update(const float* v, float* d, const float* w, float* r, const float* m, int k, long long n, double* rr,
                  double* work, hipStream_t s);
// R -= W M with rr[j] = |R_j|^2 (work: k * block_combine_blocks(n) doubles)
void block_residual_update(const float* w, float* r, const float* m, int k, long long n, double* rr, double* work,
                           hipStream_t s);
// Y0 = X0 + U M0 and Y1 = X1 + U M1 in one pass over U (k <= 48 columns in and out)
void block_combine2(const float* u, const float* x0, float* y0, const float* m0, const float* x1, float* y1,
                    const float* m1, int k, long long n, hipStream_t s);
// D += U M; columns j with flag[j] != 0 then x_j += scale[j] D_j (FP64) and D_j = 0, in the same pass
void block_combine_flush(const float* u, float* d, const float* m, int k, long long n, double* x, const double* scale,
                         const int* flag, hipStream_t s);
void block_small_product(const double* S, const double* H, int k, double scale, float* out, hipStream_t s);
// y[m] (+)= sum_r m[r][m] u[r] for k (<= 64) rows u of length n and an mcols-column device matrix m
void combine_device(const double* u, long long n, int k, const double* m, int mcols, double* y, int accumulate,
                    hipStream_t s);
// rigid-body kernel of floating fixtures (rigid.hip); coef is 6 x nk row-major over the rigid modes
// (three translations, rotations about x, y, z) about `center`
void fine_kernel(const Geom& g, const double h[3], const double center[3], const double* coef, int nk,
                 const uint8_t* mask, double* out, double* tmp, double* work, int work_blocks, double* small,
                 hipStream_t s);
void coarsest_kernel(const Geom& g, const double h[3], const double center[3], const double* coef, int nk,
                     const std::vector<int>& live, std::vector<double>& kernel, std::vector<int>& pins);
void pin_dense(double* a, int n, const int* pins_dev, int np, hipStream_t s);
void project_dense_inverse(double* a, int n, const int* pins_dev, int np, const std::vector<double>& kernel, int nk,
                           hipStream_t s);
void flat_to_grid(const Geom& pad, int ux, int uy, int uz, const double* flat, double* grid, int k, hipStream_t s);
void grid_to_flat(const Geom& pad, int ux, int uy, int uz, const double* grid, double* flat, int k, hipStream_t s);
void pad_moduli(const Geom& pad, int ux, int uy, int uz, const double* moduli, double* m64, float* m32,
                hipStream_t s);
void scale_vector(double* v, double a, long long n, hipStream_t s);
void mask_vector(const uint8_t* mask, double* v, long long nodes, int k, hipStream_t s);
int reduce_blocks(long long n);
void compliance_sensitivity(const double* u_flat, const double* f_flat, long long ndof, int k, const double* energy,
                            long long ne, const double* dmod, double factor, double* comp, double* sens,
                            double* partial, hipStream_t s);

}  // namespace dfmi
