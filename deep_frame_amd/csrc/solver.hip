// Host side of libdfmi: handle, Galerkin hierarchy set-up, V-cycle, batched
// flexible PCG and the C ABI of include/dfmi.h.
//
// Semantics follow deep_frame/topology_multigrid.py (GeometricMultigrid):
//   levels/padding :58-63, masks :138-144, :180-181, hierarchy :145-208,
//   lambda_max :214-223, Chebyshev :284-305, V-cycle :306-317,
//   precondition :318-324, PCG :407-454, solve :336-375.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>

#include "dfmi_internal.h"
#include "dfmi_learned.h"
#include "dfmi.h"

namespace dfmi {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* get_error() { return g_error.c_str(); }

// reference corner k -> corner bits q = a + 2b + 4c
static const int kRefToBits[8] = {0, 1, 3, 2, 4, 5, 7, 6};

// Device buffers are allocated and released stream-ordered (hipMallocAsync / hipFreeAsync from the device's
// memory pool, which keeps released memory for reuse) on the stream of the entry point that runs, so that
// releasing a temporary does not wait for the whole device: hipFree does, which stalls every other stream
// (concurrent handles on their own streams) and empties the queue of this one.  Every entry point of a handle
// sets that stream (StreamScope) and returns with its stream synchronised, so a buffer is never released on a
// stream other than the one its last use was queued on while that use is in flight.  Outside an entry point
// of a handle the plain allocator is used.
static thread_local hipStream_t t_stream = nullptr;
static thread_local bool t_stream_set = false;

struct StreamScope {
  hipStream_t prev;
  bool prev_set;
  explicit StreamScope(hipStream_t s) : prev(t_stream), prev_set(t_stream_set) {
    t_stream = s;
    t_stream_set = true;
  }
  ~StreamScope() {
    t_stream = prev;
    t_stream_set = prev_set;
  }
};

// keep the memory the stream-ordered allocator releases (the default threshold returns it to the system at
// every synchronisation)
static void keep_pool_memory() {
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    hipMemPool_t pool;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return;
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  });
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool pooled = false;           // allocated from the stream-ordered pool
  hipStream_t home = nullptr;    // the stream it was allocated on
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), pooled(o.pooled), home(o.home) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      pooled = o.pooled;
      home = o.home;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  void ensure(size_t n) {
    if (n <= bytes) return;
    release();
    if (t_stream_set) {
      DFMI_HIP(hipMallocAsync(&p, n, t_stream));
      pooled = true;
      home = t_stream;
    } else {
      DFMI_HIP(hipMalloc(&p, n));
      pooled = false;
    }
    bytes = n;
  }
  void release() {
    if (p) {
      if (!pooled)
        (void)hipFree(p);
      else if (t_stream_set) {
        if (home != t_stream) (void)hipStreamSynchronize(home);  // a handle whose stream was changed
        (void)hipFreeAsync(p, t_stream);
      }
      else {  // released outside an entry point of a handle: the device is idle for it only after a full wait
        (void)hipDeviceSynchronize();
        (void)hipFree(p);
      }
    }
    p = nullptr;
    bytes = 0;
    pooled = false;
  }
  template <typename T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
  ~DevBuf() { release(); }
};

struct Level {
  Geom g{};
  DevBuf mask, dinv64, dinv32, st64, st32;
  double lambda_max = 0.0;
  // V-cycle work vectors (precision of the V-cycle), capacity k_cap columns
  DevBuf rhs, x, xa, res;
  // level 1 with a Krylov coarse solve: first direction, its product, the second product (interleaved
  // vectors), the step scalars and the dot partials
  DevBuf kp, kq, kaz, kscal, kpart;
};

struct Hierarchy {
  bool built = false;
  std::vector<std::unique_ptr<Level>> levels;
  int coarsest_n = 0;
  DevBuf live_dofs, ainv, work_in, work_out, work_split;
  double setup_s = 0.0;
  // rigid-body kernel left free by the Dirichlet DOFs (floating fixtures): nk orthonormal fine-level
  // vectors, the Gram work buffer and the nk x k projection coefficients
  int nk = 0;
  DevBuf kern, kwork, kcoef;
  // |A_l B_l| / |A_l R_l| per level for slots with a free kernel (empty otherwise)
  std::vector<double> kernel_check;
  // completion of the coarsest inverse when it runs on a side stream (null: ordered on the handle's stream);
  // the coarsest solve waits for it on the device
  hipEvent_t ainv_ready = nullptr;
};

// Before a hierarchy is replaced or dropped: order the handle's stream after its pending coarsest inverse, so
// that the stream-ordered release of its buffers cannot hand them out while the side stream still writes them.
static void retire(Hierarchy& H, hipStream_t s) {
  if (!H.ainv_ready) return;
  (void)hipStreamWaitEvent(s, H.ainv_ready, 0);
  (void)hipEventDestroy(H.ainv_ready);
  H.ainv_ready = nullptr;
}

constexpr int KERNEL_GRAM_BLOCKS = 512;

struct Handle {
  int ux = 0, uy = 0, uz = 0;
  Geom pad{};
  int nlev = 0;
  double spacing[3]{};
  double ke_bits[576]{};
  KeBlocks kb{};
  dfmi_settings st{};
  hipStream_t stream = nullptr;
  DevBuf m64, m32, ke_dev, kediag_dev;
  // fine level on the matrix cores: per-lane split-bf16 Ke operands and the tile chosen per column count
  DevBuf atab;
  std::map<int, FineMfmaTile> mfma_tiles;
  // unit-density element mass in the corner-Hadamard basis (dfmi_set_mass) and a padded element field
  double mdiag[8]{};
  bool have_mass = false;
  DevBuf mass_pad;
  bool have_moduli = false;
  Hierarchy slots[8];
  // PCG workspace (padded grid, FP64)
  int k_cap = 0;
  DevBuf b, x, r, z, zn, p, ap, tmp, partial, scal, active;
  // block Krylov solve: Gram and norm partials, and the small k x k matrices (S, W^T Z, step coefficients)
  DevBuf bk_work, bk_mat;
  std::vector<Geom> geoms;
  // learned coarse correction (not owned) and training-sample harvest (caller's device buffer)
  NetImpl* net = nullptr;
  float* harvest = nullptr;
  long long harvest_cap = 0, harvest_count = 0;
  // largest relative recursive residual over the active columns after each PCG iteration of the last solve
  std::vector<double> history;
  // side streams of the hierarchy set-up (default priority): the coarse levels' power iterations and the
  // coarsest inverse run beside the fine level's power iteration
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  Handle() = default;
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  ~Handle() {
    for (Hierarchy& H : slots)
      if (H.ainv_ready) (void)hipEventSynchronize(H.ainv_ready), (void)hipEventDestroy(H.ainv_ready);
    for (hipStream_t& q : side)
      if (q) (void)hipStreamSynchronize(q), (void)hipStreamDestroy(q), q = nullptr;
  }
};

static void hex8_blocks(const double* ke_bits, KeBlocks& kb) {
  // B' = T Ke T^T with T = H (x) I3, H[m][q] = (-1)^popcount(m & q)
  double t[24][24];
  for (int m = 0; m < 8; ++m)
    for (int q = 0; q < 8; ++q)
      for (int c = 0; c < 3; ++c)
        for (int cc = 0; cc < 3; ++cc)
          t[3 * m + c][3 * q + cc] = (c == cc) ? ((__builtin_popcount(m & q) & 1) ? -1.0 : 1.0) : 0.0;
  double tmp[24][24], bp[24][24];
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      double s = 0;
      for (int k = 0; k < 24; ++k) s += t[i][k] * ke_bits[k * 24 + j];
      tmp[i][j] = s;
    }
  double big = 0;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      double s = 0;
      for (int k = 0; k < 24; ++k) s += tmp[i][k] * t[j][k];
      bp[i][j] = s;
      big = std::max(big, std::fabs(s));
    }
  // entries outside the parity classes must vanish (mirror symmetry of the brick element)
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      const int pi = (i / 3) ^ (1 << (i % 3)), pj = (j / 3) ^ (1 << (j % 3));
      if (pi != pj && std::fabs(bp[i][j]) > 1e-10 * big)
        throw std::runtime_error("element stiffness lacks the mirror symmetry of an axis-aligned brick");
    }
  for (int p = 0; p < 8; ++p)
    for (int c = 0; c < 3; ++c)
      for (int cc = 0; cc < 3; ++cc) {
        const int i = 3 * (p ^ (1 << c)) + c, j = 3 * (p ^ (1 << cc)) + cc;
        kb.b[p][c][cc] = bp[i][j] / 64.0;
      }
}

// The coarsest level is solved by an explicit dense inverse (n^2 FP64); above this many DOFs a further
// level is added instead (the reference's sparse direct coarsest solve takes coarsest_dofs up to 80000).
constexpr int DENSE_COARSEST_MAX = 40000;

static int level_count(const int shape[3], const dfmi_settings& s, int padded[3]) {
  int count = 2;
  if (s.levels > 0) {
    count = s.levels;
  } else {
    auto ok = [&](int cnt) {
      double coarse = 1.0;
      for (int a = 0; a < 3; ++a) {
        if (std::ceil(shape[a] / std::pow(2.0, cnt)) < 1) return false;
        coarse *= std::ceil(shape[a] / std::pow(2.0, cnt - 1)) + 1;
      }
      return coarse * 3 > std::min(s.coarsest_dofs, DENSE_COARSEST_MAX);
    };
    while (count < s.max_levels && ok(count)) ++count;
  }
  const int factor = 1 << (count - 1);
  for (int a = 0; a < 3; ++a) padded[a] = (shape[a] + factor - 1) / factor * factor;
  return count;
}

template <typename T>
static T* vec(Level& L, DevBuf Level::*field) {
  return (L.*field).as<T>();
}

static long long col(const Geom& g) { return g.nodes() * 3; }

// ---- V-cycle -----------------------------------------------------------------------------
struct VArgs {
  Handle* h;
  Hierarchy* H;
  int k;
  const int* active;
  // when set, the fine level's last post-smoothing step also writes per-workgroup partials of b.x (the
  // V-cycle's input dotted with its output), k x (*final_blocks) values
  double* final_dot = nullptr;
  int* final_blocks = nullptr;
};

template <typename T>
static const T* level_moduli(Handle* h);
template <>
const float* level_moduli<float>(Handle* h) {
  return h->m32.as<float>();
}
template <>
const double* level_moduli<double>(Handle* h) {
  return h->m64.as<double>();
}

template <typename T>
static const T* level_dinv(Level& L);
template <>
const float* level_dinv<float>(Level& L) {
  return L.dinv32.as<float>();
}
template <>
const double* level_dinv<double>(Level& L) {
  return L.dinv64.as<double>();
}

template <typename T>
static const T* level_stencil(Level& L);
template <>
const float* level_stencil<float>(Level& L) {
  return L.st32.as<float>();
}
template <>
const double* level_stencil<double>(Level& L) {
  return L.st64.as<double>();
}

// Workgroup tile of the matrix-core fine kernel for k columns: the autotuned one when tune_fine has run
// for k, else a default with about two thousand workgroups.
static FineMfmaTile fine_tile(Handle* h, int k) {
  auto it = h->mfma_tiles.find(k);
  if (it != h->mfma_tiles.end()) return it->second;
  const auto cands = fine_mfma_candidates(h->pad, k);
  for (const auto& c : cands)
    if (c.tye == 16) {
      const long long wg = (long long)((h->pad.ny + 14) / 15) * k * ((h->pad.nx + c.xchunk - 1) / c.xchunk);
      if (wg >= 2048) return c;
    }
  return cands.front();
}

static bool use_mfma(Handle* h, const Geom& g, int mode) {
  return h->st.precision == 1 && h->st.fine_smoother == 1 && fine_mfma_supported(g, mode);
}

template <typename T>
static void level_apply(VArgs& v, int idx, int mode, ApplyArgs& a, int* nblocks = nullptr) {
  Level& L = *v.H->levels[idx];
  a.k = v.k;
  a.active = v.active;
  if (idx == 0 && sizeof(T) == sizeof(float) && use_mfma(v.h, L.g, mode))
    fine_apply_mfma(L.g, v.h->m32.as<float>(), L.mask.as<uint8_t>(), v.h->atab.as<unsigned>(), mode, a,
                    fine_tile(v.h, v.k), v.h->stream);
  else if (idx == 0) {
    a.premasked = 1;  // V-cycle iterates are masked outputs of the operator kernels and the prolongation
    fine_apply<T>(L.g, level_moduli<T>(v.h), L.mask.as<uint8_t>(), v.h->kb, mode, a, v.h->stream, nblocks);
  }
  else
    stencil_apply<T>(L.g, level_stencil<T>(L), L.mask.as<uint8_t>(), mode, a, interleaved(v.k), v.h->stream, nullptr);
}

// vector layout of a level inside the V-cycle: column-major on the fine level, interleaved below
static VLay level_layout(const Level& L, int idx, int k) { return idx == 0 ? column_major(L.g) : interleaved(k); }

// Requests the b.x partials on the fine level's last post-smoothing step when the caller asked for them
// (VArgs::final_dot) and the step runs on the vector kernel, which forms them.
static void final_dot(VArgs& v, int idx, bool last, ApplyArgs& a) {
  if (!v.final_dot || idx != 0 || !last || use_mfma(v.h, v.H->levels[0]->g, MODE_CHEB)) return;
  a.partial = v.final_dot;
}

// Step coefficients of the Chebyshev smoother (topology_multigrid.py:284-305 in the three-term form) for `count` steps: the first step is x1 = x0 + c2[0] Dinv r0, step
// j >= 1 is x_{j+1} = x_j + c1[j] (x_j - x_{j-1}) + c2[j] Dinv (b - A x_j).
//   kind 1 (the reference's polynomial, topology_multigrid.py:291-304): bounds [upper / ratio, upper], theta and
//     delta their centre and half-width, sigma = theta / delta, rho_0 = 1 / sigma, c2[0] = 1 / theta and
//     rho_j = 1 / (2 sigma - rho_{j-1}), c1[j] = rho_j rho_{j-1}, c2[j] = 2 rho_j / delta;
//   kind 4 (Chebyshev polynomials of the fourth kind, Lottes 2023): c2[0] = 4 / (3 upper) and
//     c1[j] = (2j - 1) / (2j + 3), c2[j] = (8j + 4) / ((2j + 3) upper); the error polynomial after k steps is
//     W_k(1 - 2t) / (2k + 1), t = eigenvalue / upper, W_k the fourth-kind Chebyshev polynomial.
// upper = 1.1 lambda_max in both.
static void chebyshev_coefficients(int kind, double lambda_max, double ratio, int count, std::vector<double>& c1,
                                   std::vector<double>& c2) {
  const double upper = 1.1 * lambda_max;
  c1.assign(std::max(count, 1), 0.0);
  c2.assign(std::max(count, 1), 0.0);
  if (kind == 4) {
    c2[0] = 4.0 / (3.0 * upper);
    for (int j = 1; j < count; ++j) {
      c1[j] = (2.0 * j - 1.0) / (2.0 * j + 3.0);
      c2[j] = (8.0 * j + 4.0) / ((2.0 * j + 3.0) * upper);
    }
    return;
  }
  const double lower = upper / ratio;
  const double theta = 0.5 * (upper + lower), delta = 0.5 * (upper - lower);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  c2[0] = 1.0 / theta;
  for (int j = 1; j < count; ++j) {
    const double nr = 1.0 / (2.0 * sigma - rho);
    c1[j] = nr * rho;
    c2[j] = 2.0 * nr / delta;
    rho = nr;
  }
}

// Chebyshev smoothing of A x = b on level idx (coefficients above), from x = 0 or from the level's x.  On exit
// the iterate is in L.x.
template <typename T>
static void chebyshev(VArgs& v, int idx, const T* b, bool from_zero, int count) {
  Level& L = *v.H->levels[idx];
  std::vector<double> c1, c2;
  chebyshev_coefficients(v.h->st.chebyshev_kind, L.lambda_max, v.h->st.chebyshev_ratio, count, c1, c2);
  T* cur = L.x.as<T>();
  T* other = L.xa.as<T>();
  const T* dinv = level_dinv<T>(L);
  bool prev_zero;
  int j0 = 1;
  if (from_zero && idx == 0 && count >= 2) {
    // fine level: x1 = c2[0] Dinv b is formed inside the first step (no separate pass) and re-formed from b as
    // the previous iterate of the second step
    ApplyArgs a{};
    a.b = b;
    a.dinv = dinv;
    a.c0 = c2[0];
    a.c1 = c1[1];
    a.c2 = c2[1];
    a.y = cur;
    level_apply<T>(v, idx, MODE_CHEB0, a);
    if (count >= 3) {
      ApplyArgs a2{};
      a2.x = cur;
      a2.y = other;
      a2.b = b;
      a2.dinv = dinv;
      a2.c0 = c2[0];
      a2.c1 = c1[2];
      a2.c2 = c2[2];
      level_apply<T>(v, idx, MODE_CHEB1, a2);
      std::swap(cur, other);
    }
    prev_zero = false;
    j0 = std::min(count, 3);
  } else if (from_zero) {
    cheb_start<T>(b, dinv, c2[0], cur, L.g.nodes(), v.k, v.active, level_layout(L, idx, v.k), v.h->stream);
    prev_zero = true;
  } else {
    ApplyArgs a{};
    a.x = cur;
    a.y = other;
    a.b = b;
    a.xp = nullptr;
    a.dinv = dinv;
    a.c1 = 0.0;
    a.c2 = c2[0];
    final_dot(v, idx, count == 1, a);
    level_apply<T>(v, idx, MODE_CHEB, a, v.final_blocks);
    std::swap(cur, other);
    prev_zero = false;
  }
  for (int j = j0; j < count; ++j) {
    ApplyArgs a{};
    a.x = cur;
    a.y = other;
    a.b = b;
    a.xp = prev_zero ? nullptr : other;
    a.dinv = dinv;
    a.c1 = c1[j];
    a.c2 = c2[j];
    final_dot(v, idx, !from_zero && j == count - 1, a);
    level_apply<T>(v, idx, MODE_CHEB, a, v.final_blocks);
    std::swap(cur, other);
    prev_zero = false;
  }
  if (cur != L.x.as<T>()) std::swap(L.x.p, L.xa.p), std::swap(L.x.bytes, L.xa.bytes);
}

// Damped Jacobi smoothing of A x = b (topology_multigrid.py:286-290): x_{j+1} = x_j + w Dinv (b - A x_j) with
// w = damping / lambda_max, from x_0 = 0 (first sweep x_1 = w Dinv b) or from the level's x.  Each sweep is
// the Chebyshev step form with c1 = 0 and c2 = w, so it runs on the same fused operator kernels (on the fine
// level the first two sweeps from zero are one launch, as for Chebyshev).  On exit the iterate is in L.x.
template <typename T>
static void jacobi(VArgs& v, int idx, const T* b, bool from_zero, int count) {
  Level& L = *v.H->levels[idx];
  const double w = v.h->st.damping / L.lambda_max;
  T* cur = L.x.as<T>();
  T* other = L.xa.as<T>();
  const T* dinv = level_dinv<T>(L);
  int j0 = 0;
  if (from_zero && idx == 0 && count >= 2) {
    ApplyArgs a{};
    a.b = b;
    a.dinv = dinv;
    a.c0 = w;
    a.c1 = 0.0;
    a.c2 = w;
    a.y = cur;
    level_apply<T>(v, idx, MODE_CHEB0, a);
    j0 = 2;
    if (count >= 3) {
      ApplyArgs a2{};
      a2.x = cur;
      a2.y = other;
      a2.b = b;
      a2.dinv = dinv;
      a2.c0 = w;
      a2.c1 = 0.0;
      a2.c2 = w;
      level_apply<T>(v, idx, MODE_CHEB1, a2);
      std::swap(cur, other);
      j0 = 3;
    }
  } else if (from_zero) {
    cheb_start<T>(b, dinv, w, cur, L.g.nodes(), v.k, v.active, level_layout(L, idx, v.k), v.h->stream);
    j0 = 1;
  }
  for (int j = j0; j < count; ++j) {
    ApplyArgs a{};
    a.x = cur;
    a.y = other;
    a.b = b;
    a.xp = nullptr;
    a.dinv = dinv;
    a.c1 = 0.0;
    a.c2 = w;
    final_dot(v, idx, !from_zero && j == count - 1, a);
    level_apply<T>(v, idx, MODE_CHEB, a, v.final_blocks);
    std::swap(cur, other);
  }
  if (cur != L.x.as<T>()) std::swap(L.x.p, L.xa.p), std::swap(L.x.bytes, L.xa.bytes);
}

// Smoother steps on level idx (topology_multigrid.py:284-285): round(degree * growth^idx) Chebyshev steps or
// round(sweeps * growth^idx) Jacobi sweeps, at least one.
static int smoother_count(const Handle* h, int idx) {
  const double base = h->st.smoother == 1 ? (double)h->st.sweeps : (double)h->st.chebyshev_degree;
  // ties to even, as Python's round() in the reference
  return std::max(1, (int)std::nearbyint(base * std::pow(h->st.growth, (double)idx)));
}

template <typename T>
static void smooth(VArgs& v, int idx, const T* b, bool from_zero) {
  const int count = smoother_count(v.h, idx);
  if (v.h->st.smoother == 1)
    jacobi<T>(v, idx, b, from_zero, count);
  else
    chebyshev<T>(v, idx, b, from_zero, count);
}

// Learned coarse correction on the fine level, between the coarse-grid correction and the
// post-smoother: g = R (b - A x) on level 1, delta = Dinv g + Dinv^(1/2) s N(Dinv^(1/2) g / s, f), scaled per column by the
// energy-optimal step max(0, delta.g / delta.A_1 delta), then x += P delta.  With the Galerkin level-1
// operator A_1 = P^T A P this step never increases the A-norm error of x.  The same residual g is
// what the training harvest records.
template <typename T>
static void learned_correction(VArgs& v, const T* b) {
  if constexpr (sizeof(T) != sizeof(float)) {
    if (v.h->net) throw std::invalid_argument("the learned correction needs the fp32 V-cycle precision");
    return;
  } else {
    Hierarchy& H = *v.H;
    if (H.levels.size() < 2) return;
    Level& L = *H.levels[0];
    Level& C = *H.levels[1];
    hipStream_t s = v.h->stream;
    const VLay cl = interleaved(v.k);
    ApplyArgs a{};
    a.x = L.x.as<float>();
    a.y = L.res.as<float>();
    a.b = b;
    level_apply<float>(v, 0, MODE_RESID, a);
    restrict_full<float>(L.g, L.res.as<float>(), column_major(L.g), C.g, C.mask.as<uint8_t>(), C.rhs.as<float>(), cl,
                         v.k, v.active, s);
    const long long nodes = C.g.nodes();
    if (v.h->harvest && v.h->harvest_count + v.k <= v.h->harvest_cap) {
      gather_samples(C.rhs.as<float>(), cl, nodes, v.k, v.active, v.h->harvest + v.h->harvest_count * 3 * nodes, s);
      v.h->harvest_count += v.k;
    }
    if (!v.h->net) return;
    net_set_stream(v.h->net, s);
    net_correct(v.h->net, C.g, C.rhs.as<float>(), cl, v.k, C.dinv32.as<float>(), v.active, C.x.as<float>());
    ApplyArgs c{};
    c.x = C.x.as<float>();
    c.y = C.xa.as<float>();
    level_apply<float>(v, 1, MODE_APPLY, c);
    net_energy_step(v.h->net, C.rhs.as<float>(), C.xa.as<float>(), C.x.as<float>(), cl, nodes, v.k, v.active);
    prolong_add<float>(C.g, C.x.as<float>(), cl, L.g, L.mask.as<uint8_t>(), L.x.as<float>(), column_major(L.g), v.k,
                       v.active, s);
  }
}

template <typename T>
static void vcycle(VArgs& v, int idx, const T* b);

// Solve of a coarse-grid correction on level lv >= 1 by m = coarse_krylov flexible-CG steps, each preconditioned
// by the V-cycle from level lv (a K-cycle; the V-cycle there uses the same Krylov solve one level further down
// while lv < krylov_levels): z0 = V(g), q = A z0, alpha0 = g.z0 / z0.q; for m = 2 then r1 = g - alpha0 q,
// z1 = V(r1), beta = (r1.z1 - r1.z0) / g.z0 and p1 = z1 + beta z0, whose product A p1 = A z1 + beta q needs
// one more operator apply, and x = alpha0 z0 + alpha1 p1 with alpha1 = r1.z1 / p1.A p1.  g (level lv's rhs
// buffer) is overwritten by r1; the result is left in level lv's x, as the V-cycle leaves it.
template <typename T>
static void krylov_coarse(VArgs& v, int lv, T* g) {
  Level& C = *v.H->levels[lv];
  hipStream_t s = v.h->stream;
  const int k = v.k, m = v.h->st.coarse_krylov;
  const long long nodes = C.g.nodes();
  double* ks = C.kscal.as<double>();
  double* part = C.kpart.as<double>();
  const int nb = level_dot_blocks(nodes);
  vcycle<T>(v, lv, g);
  std::swap(C.x.p, C.kp.p), std::swap(C.x.bytes, C.kp.bytes);  // kp = z0
  const T* z0 = C.kp.as<T>();
  ApplyArgs a{};
  a.x = z0;
  a.y = C.kq.as<T>();
  level_apply<T>(v, lv, MODE_APPLY, a);
  LevelPairs<T> d1{{g, z0, nullptr, nullptr}, {z0, C.kq.as<T>(), nullptr, nullptr}};
  level_dots<T>(d1, 2, nodes, k, v.active, part, nb, s);
  reduce_partials(part, nb, 2 * k, ks, s);
  krylov_scalars(ks, k, v.active, 1, s);
  if (m >= 2) {
    level_axpy<T>(g, C.kq.as<T>(), ks + 6 * (size_t)k, nodes, k, v.active, s);  // r1
    vcycle<T>(v, lv, g);                                                        // z1 in C.x
    ApplyArgs b{};
    b.x = C.x.as<T>();
    b.y = C.kaz.as<T>();
    level_apply<T>(v, lv, MODE_APPLY, b);
    const T* z1 = C.x.as<T>();
    LevelPairs<T> d2{{g, g, z1, z1}, {z1, z0, C.kaz.as<T>(), C.kq.as<T>()}};
    level_dots<T>(d2, 4, nodes, k, v.active, part, nb, s);
    reduce_partials(part, nb, 4 * k, ks + 2 * (size_t)k, s);
    krylov_scalars(ks, k, v.active, 2, s);
  }
  level_combine<T>(C.x.as<T>(), z0, ks + 7 * (size_t)k, ks + 8 * (size_t)k, nodes, k, v.active, s);
}

// Whether the coarse-grid correction of level idx is the Krylov solve on level idx + 1: coarse_krylov steps
// on the first krylov_levels levels whose coarser level is not the coarsest (slots with a free rigid kernel keep
// the V-cycle: their coarse operators are singular).
static bool krylov_at(const VArgs& v, int idx) {
  return v.h->st.coarse_krylov > 0 && idx < v.h->st.krylov_levels && idx + 2 < (int)v.H->levels.size() &&
         v.H->nk == 0;
}

// V-cycle on level idx for right-hand side b (device, T); result in levels[idx].x
template <typename T>
static void vcycle(VArgs& v, int idx, const T* b) {
  Hierarchy& H = *v.H;
  Level& L = *H.levels[idx];
  hipStream_t s = v.h->stream;
  if (idx == (int)H.levels.size() - 1) {
    if (H.ainv_ready) DFMI_HIP(hipStreamWaitEvent(s, H.ainv_ready, 0));
    coarsest_solve<T>(L.g, H.ainv.as<double>(), H.live_dofs.as<int>(), H.coarsest_n, b, L.x.as<T>(),
                      level_layout(L, idx, v.k), H.work_in.as<double>(), H.work_out.as<double>(),
                      H.work_split.as<double>(), v.k, v.active, s);
    return;
  }
  Level& C = *H.levels[idx + 1];
  smooth<T>(v, idx, b, true);
  // Between the column-major fine level and the interleaved level 1 the transfers run column-major on
  // both sides (coalesced), with a tiled layout change of the level-1 vector through its free res buffer.
  const bool change = idx == 0 && v.k > 1;
  // W-cycle: a second coarse-grid correction, except above the coarsest level (topology_multigrid.py:312)
  const int corrections = (v.h->st.cycle == 1 && idx + 2 < (int)H.levels.size()) ? 2 : 1;
  for (int rep = 0; rep < corrections; ++rep) {
    ApplyArgs a{};
    a.x = L.x.as<T>();
    a.y = L.res.as<T>();
    a.b = b;
    level_apply<T>(v, idx, MODE_RESID, a);
    if (change) {
      restrict_full<T>(L.g, L.res.as<T>(), column_major(L.g), C.g, C.mask.as<uint8_t>(), C.res.as<T>(),
                       column_major(C.g), v.k, v.active, s);
      relayout<T>(C.res.as<T>(), C.rhs.as<T>(), C.g.nodes(), v.k, false, v.active, s);
    } else {
      restrict_full<T>(L.g, L.res.as<T>(), level_layout(L, idx, v.k), C.g, C.mask.as<uint8_t>(), C.rhs.as<T>(),
                       level_layout(C, idx + 1, v.k), v.k, v.active, s);
    }
    if (krylov_at(v, idx))
      krylov_coarse<T>(v, idx + 1, C.rhs.as<T>());
    else
      vcycle<T>(v, idx + 1, C.rhs.as<T>());
    if (change) {
      relayout<T>(C.x.as<T>(), C.res.as<T>(), C.g.nodes(), v.k, true, v.active, s);
      prolong_add<T>(C.g, C.res.as<T>(), column_major(C.g), L.g, L.mask.as<uint8_t>(), L.x.as<T>(), column_major(L.g),
                     v.k, v.active, s);
    } else {
      prolong_add<T>(C.g, C.x.as<T>(), level_layout(C, idx + 1, v.k), L.g, L.mask.as<uint8_t>(), L.x.as<T>(),
                     level_layout(L, idx, v.k), v.k, v.active, s);
    }
  }
  if (idx == 0 && (v.h->net || v.h->harvest)) learned_correction<T>(v, b);
  smooth<T>(v, idx, b, false);
}

// ---- hierarchy ------------------------------------------------------------------------------
static void sync(Handle* h) { DFMI_HIP(hipStreamSynchronize(h->stream)); }

static void ensure_level_vectors(Handle* h, Hierarchy& H, int k) {
  const size_t el = h->st.precision == 1 ? sizeof(float) : sizeof(double);
  for (auto& Lp : H.levels) {
    Level& L = *Lp;
    const size_t n = (size_t)k * col(L.g) * el;
    L.rhs.ensure(n);
    L.x.ensure(n);
    L.xa.ensure(n);
    L.res.ensure(n);
  }
  for (int l = 1; h->st.coarse_krylov > 0 && l <= h->st.krylov_levels && l + 1 < (int)H.levels.size(); ++l) {
    Level& C = *H.levels[l];
    const size_t n = (size_t)k * col(C.g) * el;
    C.kp.ensure(n);
    C.kq.ensure(n);
    C.kaz.ensure(n);
    C.kscal.ensure(sizeof(double) * 9 * (size_t)k);
    C.kpart.ensure(sizeof(double) * 4 * (size_t)k * level_dot_blocks(C.g.nodes()));
  }
  if (H.coarsest_n > 0) {
    H.work_in.ensure(sizeof(double) * (size_t)k * H.coarsest_n);
    H.work_out.ensure(sizeof(double) * (size_t)k * H.coarsest_n);
    H.work_split.ensure(sizeof(double) * (size_t)coarsest_splits() * k * H.coarsest_n);
  }
}

// Autotuner of the matrix-core fine kernel: times every candidate tile (MODE_APPLY on the fine level's
// x and res vectors, which hold nothing between V-cycles) with HIP events on the handle's stream and
// keeps the fastest for this column count.  Runs once per k.
static FineMfmaTile tune_fine(Handle* h, Hierarchy& H, int k, double* best_ms = nullptr) {
  Level& F = *H.levels[0];
  auto found = h->mfma_tiles.find(k);
  if (found != h->mfma_tiles.end() && !best_ms) return found->second;
  const long long n = (long long)k * col(F.g);
  fill_zero<float>(F.x.as<float>(), n, h->stream);
  ApplyArgs a{};
  a.x = F.x.as<float>();
  a.y = F.res.as<float>();
  a.k = k;
  hipEvent_t e0, e1;
  DFMI_HIP(hipEventCreate(&e0));
  DFMI_HIP(hipEventCreate(&e1));
  FineMfmaTile best{};
  float best_t = -1.f;
  try {
    for (const auto& c : fine_mfma_candidates(F.g, k)) {
      fine_apply_mfma(F.g, h->m32.as<float>(), F.mask.as<uint8_t>(), h->atab.as<unsigned>(), MODE_APPLY, a, c,
                      h->stream);
      DFMI_HIP(hipEventRecord(e0, h->stream));
      for (int rep = 0; rep < 3; ++rep)
        fine_apply_mfma(F.g, h->m32.as<float>(), F.mask.as<uint8_t>(), h->atab.as<unsigned>(), MODE_APPLY, a, c,
                        h->stream);
      DFMI_HIP(hipEventRecord(e1, h->stream));
      DFMI_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      DFMI_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (best_t < 0 || ms < best_t) best_t = ms, best = c;
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->mfma_tiles[k] = best;
  if (best_ms) *best_ms = best_t / 3.0;
  return best;
}

static void maybe_tune_fine(Handle* h, Hierarchy& H, int k) {
  if (use_mfma(h, H.levels[0]->g, MODE_CHEB)) tune_fine(h, H, k);
}

static void ensure_workspace(Handle* h, int k) {
  const size_t n = sizeof(double) * (size_t)k * col(h->pad);
  for (DevBuf* b : {&h->b, &h->x, &h->r, &h->z, &h->zn, &h->p, &h->ap, &h->tmp}) b->ensure(n);
  const int nb = std::max(fine_blocks(h->pad), 2 * reduce_blocks(col(h->pad)));
  h->partial.ensure(sizeof(double) * (size_t)k * nb);
  h->scal.ensure(sizeof(double) * (size_t)k * 16);
  h->active.ensure(sizeof(int) * (size_t)k);
  h->k_cap = std::max(h->k_cap, k);
}

__global__ void k_to_float(const double* a, float* b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_inverse_diag(const double* d, const uint8_t* mask, double* inv64, float* inv32, long long nodes) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes * 3) return;
  const bool on = (mask[i / 3] >> (i % 3)) & 1;
  const double v = on && d[i] > 0 ? 1.0 / d[i] : 0.0;
  inv64[i] = v;
  inv32[i] = (float)v;
}

// Rediscretised coarse level (topology_multigrid.py:160-166): element field of the coarse grid g as the mean
// of its 2x2x2 children on the finer grid f (avg_pool3d), times `scale`.
__global__ void k_pool_moduli(Geom f, Geom g, const double* __restrict__ fine, double scale, double* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.elems()) return;
  const int l = (int)(e % g.ez), j = (int)((e / g.ez) % g.ey), i = (int)(e / ((long long)g.ez * g.ey));
  double acc = 0.0;
  for (int q = 0; q < 8; ++q)
    acc += fine[((long long)(2 * i + (q & 1)) * f.ey + 2 * j + ((q >> 1) & 1)) * f.ez + 2 * l + ((q >> 2) & 1)];
  out[e] = scale * 0.125 * acc;
}

// mask of coarse node (i, j, l) = mask of finer node (2i, 2j, 2l) (the injection of the finer level's mask)
__global__ void k_inject_mask(Geom f, Geom g, const uint8_t* __restrict__ fine, uint8_t* __restrict__ out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nodes()) return;
  const int l = (int)(n % g.nz), j = (int)((n / g.nz) % g.ny), i = (int)(n / ((long long)g.nz * g.ny));
  out[n] = fine[((long long)(2 * i) * f.ny + 2 * j) * f.nz + 2 * l];
}

// v = u (one uniform [0, 1) value per DOF from a counter hash) on live DOFs, 0 elsewhere; k columns
__global__ void k_uniform_masked(const uint8_t* __restrict__ mask, double* __restrict__ v, long long nodes, int k,
                                 unsigned seed) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes * 3 * k) return;
  const long long dof = i % (nodes * 3);
  unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ULL + seed;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  v[i] = ((mask[dof / 3] >> (dof % 3)) & 1) ? u : 0.0;
}

__global__ void k_to_double(const float* a, double* b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

static void require_workspace(Handle* h) {
  if (!h->partial.p || !h->scal.p) throw std::logic_error("reduction workspace used before allocation");
}

static double column_norm(Handle* h, const double* v, long long n) {
  require_workspace(h);
  const int nb = reduce_blocks(n);
  column_dot<double>(v, v, n, 1, h->partial.as<double>(), nb, nullptr, h->stream);
  reduce_partials(h->partial.as<double>(), nb, 1, h->scal.as<double>(), h->stream);
  double out = 0;
  DFMI_HIP(hipMemcpyAsync(&out, h->scal.as<double>(), sizeof(double), hipMemcpyDeviceToHost, h->stream));
  sync(h);
  return std::sqrt(out);
}

static double device_max(Handle* h, const double* v, long long n) {
  require_workspace(h);
  const int nb = reduce_blocks(n);
  column_absmax<double>(v, n, 1, h->partial.as<double>(), nb, nullptr, h->stream);
  reduce_max(h->partial.as<double>(), nb, 1, h->scal.as<double>(), h->stream);
  double out = 0;
  DFMI_HIP(hipMemcpyAsync(&out, h->scal.as<double>(), sizeof(double), hipMemcpyDeviceToHost, h->stream));
  sync(h);
  return out;
}

// lambda_max(Dinv A) by power iteration (topology_multigrid.py:214-223), in FP64, queued on stream `st` with the
// caller's buffers (v, y: one level vector each; nrm: 2 doubles, |w|^2 of the last step left in nrm[0]; part:
// reduce_blocks partials).  The iterate is normalised on the device (w = Dinv A v, v = w / |w|), so nothing waits
// for the host; lambda_max = sqrt(nrm[0]) once the stream has run.
struct PowerBufs {
  DevBuf v, y, nrm, part;
};

static void spectrum_async(Handle* h, Hierarchy& H, int idx, hipStream_t st, PowerBufs& pb) {
  Level& L = *H.levels[idx];
  const long long n = col(L.g);
  double* sq = pb.nrm.as<double>();       // |w|^2
  double* inv = pb.nrm.as<double>() + 1;  // 1 / |w|
  // no host-to-device copy here: a pageable copy would wait for the other streams of the set-up
  power_init(L.mask.as<uint8_t>(), pb.v.as<double>(), L.g.nodes(), (unsigned)h->st.seed + 7919u * idx, st);
  const int nb = reduce_blocks(n);
  double* part = pb.part.as<double>();
  for (int it = 0; it < h->st.power_iterations; ++it) {
    ApplyArgs a{};
    a.x = pb.v.as<double>();
    a.y = pb.y.as<double>();
    a.k = 1;
    if (idx == 0)
      fine_apply<double>(L.g, h->m64.as<double>(), L.mask.as<uint8_t>(), h->kb, MODE_APPLY, a, st, nullptr);
    else
      stencil_apply<double>(L.g, L.st64.as<double>(), L.mask.as<uint8_t>(), MODE_APPLY, a, column_major(L.g), st,
                            nullptr);
    dinv_scale(pb.y.as<double>(), L.dinv64.as<double>(), nullptr, pb.y.as<double>(), L.g.nodes(), st);
    column_dot<double>(pb.y.as<double>(), pb.y.as<double>(), n, 1, part, nb, nullptr, st);
    reduce_partials(part, nb, 1, sq, st);
    inv_sqrt(sq, inv, st);
    dinv_scale(pb.y.as<double>(), nullptr, inv, pb.v.as<double>(), L.g.nodes(), st);
  }
}

// Euclidean norm of n contiguous FP64 values (set-up diagnostics; own partial buffers)
static double vector_norm(Handle* h, const double* v, long long n) {
  const int nb = reduce_blocks(n);
  DevBuf part, out;
  part.ensure(sizeof(double) * nb);
  out.ensure(sizeof(double));
  column_dot<double>(v, v, n, 1, part.as<double>(), nb, nullptr, h->stream);
  reduce_partials(part.as<double>(), nb, 1, out.as<double>(), h->stream);
  double sq = 0.0;
  DFMI_HIP(hipMemcpyAsync(&sq, out.p, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  sync(h);
  return std::sqrt(sq);
}

// hierarchy["kernel_check"] (topology_multigrid.py:325-335): per level |A_l B_l| / |A_l R_l| with B_l the nk
// orthonormal masked rigid modes of the level (about the design-grid centre, level spacing) and R_l uniform
// [0, 1) values on the level's live DOFs, both products in the V-cycle precision.
static void kernel_check(Handle* h, Hierarchy& H, const double* coef, int nk, const double center[3]) {
  H.kernel_check.clear();
  hipStream_t s = h->stream;
  const bool fp32 = h->st.precision == 1;
  for (int l = 0; l < (int)H.levels.size(); ++l) {
    Level& L = *H.levels[l];
    const long long n = col(L.g), kn = (long long)nk * n;
    const double hl[3] = {h->spacing[0] * (1 << l), h->spacing[1] * (1 << l), h->spacing[2] * (1 << l)};
    DevBuf basis, tmp, work, small, xs, ys, out;
    basis.ensure(sizeof(double) * kn);
    tmp.ensure(sizeof(double) * kn);
    work.ensure(sizeof(double) * (size_t)KERNEL_GRAM_BLOCKS * nk * 64);
    small.ensure(sizeof(double) * (size_t)nk * 64);
    out.ensure(sizeof(double) * kn);
    fine_kernel(L.g, hl, center, coef, nk, L.mask.as<uint8_t>(), basis.as<double>(), tmp.as<double>(),
                work.as<double>(), KERNEL_GRAM_BLOCKS, small.as<double>(), s);
    hipLaunchKernelGGL(k_uniform_masked, dim3((unsigned)((kn + 255) / 256)), dim3(256), 0, s, L.mask.as<uint8_t>(),
                       tmp.as<double>(), L.g.nodes(), nk, (unsigned)(h->st.seed + 104729u * (l + 1)));
    DFMI_HIP(hipGetLastError());
    double norms[2];
    for (int which = 0; which < 2; ++which) {
      const double* x = which == 0 ? basis.as<double>() : tmp.as<double>();
      ApplyArgs a{};
      a.k = nk;
      if (fp32) {
        xs.ensure(sizeof(float) * kn);
        ys.ensure(sizeof(float) * kn);
        hipLaunchKernelGGL(k_to_float, dim3((unsigned)((kn + 255) / 256)), dim3(256), 0, s, x, xs.as<float>(), kn);
        a.x = xs.as<float>();
        a.y = ys.as<float>();
        if (l == 0)
          fine_apply<float>(L.g, h->m32.as<float>(), L.mask.as<uint8_t>(), h->kb, MODE_APPLY, a, s, nullptr);
        else
          stencil_apply<float>(L.g, L.st32.as<float>(), L.mask.as<uint8_t>(), MODE_APPLY, a, column_major(L.g), s,
                               nullptr);
        hipLaunchKernelGGL(k_to_double, dim3((unsigned)((kn + 255) / 256)), dim3(256), 0, s, ys.as<float>(),
                           out.as<double>(), kn);
      } else {
        a.x = x;
        a.y = out.as<double>();
        if (l == 0)
          fine_apply<double>(L.g, h->m64.as<double>(), L.mask.as<uint8_t>(), h->kb, MODE_APPLY, a, s, nullptr);
        else
          stencil_apply<double>(L.g, L.st64.as<double>(), L.mask.as<uint8_t>(), MODE_APPLY, a, column_major(L.g), s,
                                nullptr);
      }
      DFMI_HIP(hipGetLastError());
      norms[which] = vector_norm(h, out.as<double>(), kn);
    }
    H.kernel_check.push_back(norms[0] / std::max(norms[1], 1e-300));
  }
}

static void build(Handle* h, int slot, const uint8_t* fixed_dev, const double* coef = nullptr, int nk = 0) {
  if (!h->have_moduli) throw std::runtime_error("update(moduli) must precede hierarchy construction");
  auto t0 = std::chrono::steady_clock::now();
  Hierarchy& H = h->slots[slot];
  hipStream_t s = h->stream;
  retire(H, s);
  H = Hierarchy();
  ensure_workspace(h, 1);  // reductions below use the partial-sum and scalar buffers
  for (int l = 0; l < h->nlev; ++l) {
    auto L = std::make_unique<Level>();
    L->g = h->geoms[l];
    const long long nn = L->g.nodes();
    L->mask.ensure(nn);
    L->dinv64.ensure(sizeof(double) * nn * 3);
    L->dinv32.ensure(sizeof(float) * nn * 3);
    H.levels.push_back(std::move(L));
  }
  Level& F = *H.levels[0];
  // fine mask and diagonal
  active_mask_flat(F.g, h->ux, h->uy, h->uz, h->m64.as<double>(), fixed_dev, F.mask.as<uint8_t>(), s);
  DevBuf diag;
  diag.ensure(sizeof(double) * F.g.nodes() * 3);
  fine_diagonal(F.g, h->m64.as<double>(), F.mask.as<uint8_t>(), h->kediag_dev.as<double>(), diag.as<double>(), s);
  const long long n0 = F.g.nodes() * 3;
  hipLaunchKernelGGL(k_inverse_diag, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s, diag.as<double>(),
                     F.mask.as<uint8_t>(), F.dinv64.as<double>(), F.dinv32.as<float>(), F.g.nodes());
  // rigid kernel of a floating fixture: masked modes about the centre of the design grid, orthonormal
  const double center[3] = {0.5 * h->ux * h->spacing[0], 0.5 * h->uy * h->spacing[1], 0.5 * h->uz * h->spacing[2]};
  if (nk > 0) {
    if (nk > 6 || !coef) throw std::invalid_argument("rigid kernel needs 1 to 6 coefficient columns");
    H.nk = nk;
    H.kern.ensure(sizeof(double) * (size_t)nk * n0);
    H.kwork.ensure(sizeof(double) * (size_t)KERNEL_GRAM_BLOCKS * nk * 64);
    H.kcoef.ensure(sizeof(double) * (size_t)nk * 64);
    DevBuf tmp;
    tmp.ensure(sizeof(double) * (size_t)nk * n0);
    fine_kernel(F.g, h->spacing, center, coef, nk, F.mask.as<uint8_t>(), H.kern.as<double>(), tmp.as<double>(),
                H.kwork.as<double>(), KERNEL_GRAM_BLOCKS, H.kcoef.as<double>(), s);
  }
  // Galerkin levels (P^T A P), or rediscretised ones (coarse = 1): the operator assembled on the coarse grid
  // from 2x2x2-averaged moduli with Ke scaled by the spacing ratio 2^l (3-D elasticity: Ke(2h) = 2 Ke(h)) and
  // the finer level's mask injected (topology_multigrid.py:160-166)
  const bool rediscretize = h->st.coarse == 1;
  DevBuf fst, pooled[2];
  if (!rediscretize) {
    fst.ensure(sizeof(double) * 243 * (size_t)F.g.nodes());
    fine_stencil(F.g, h->m64.as<double>(), F.mask.as<uint8_t>(), h->ke_dev.as<double>(), fst.as<double>(), s);
  }
  const double* prev = fst.as<double>();
  const double* prev_moduli = h->m64.as<double>();
  for (int l = 1; l < h->nlev; ++l) {
    Level& P = *H.levels[l - 1];
    Level& L = *H.levels[l];
    const long long nn = L.g.nodes();
    L.st64.ensure(sizeof(double) * 243 * (size_t)nn);
    if (rediscretize) {
      DevBuf& mean = pooled[l & 1];
      DevBuf scaled;
      mean.ensure(sizeof(double) * (size_t)L.g.elems());
      scaled.ensure(sizeof(double) * (size_t)L.g.elems());
      const unsigned ge = (unsigned)((L.g.elems() + 255) / 256);
      hipLaunchKernelGGL(k_pool_moduli, dim3(ge), dim3(256), 0, s, P.g, L.g, prev_moduli, 1.0, mean.as<double>());
      hipLaunchKernelGGL(k_pool_moduli, dim3(ge), dim3(256), 0, s, P.g, L.g, prev_moduli, (double)(1 << l),
                         scaled.as<double>());
      hipLaunchKernelGGL(k_inject_mask, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, P.g, L.g,
                         P.mask.as<uint8_t>(), L.mask.as<uint8_t>());
      DFMI_HIP(hipGetLastError());
      fine_stencil(L.g, scaled.as<double>(), L.mask.as<uint8_t>(), h->ke_dev.as<double>(), L.st64.as<double>(), s);
      sync(h);  // `scaled` and the older pooled field are released below
      prev_moduli = mean.as<double>();
    } else {
      stencil_rap(P.g, prev, L.g, L.st64.as<double>(), s);
    }
    DevBuf d;
    d.ensure(sizeof(double) * nn * 3);
    stencil_diagonal(L.g, L.st64.as<double>(), d.as<double>(), s);
    const double dmax = device_max(h, d.as<double>(), nn * 3);
    dead_mask(L.g, d.as<double>(), h->st.dead_ratio * dmax, L.mask.as<uint8_t>(), s);
    stencil_mask(L.g, L.st64.as<double>(), L.mask.as<uint8_t>(), s);
    hipLaunchKernelGGL(k_inverse_diag, dim3((unsigned)((nn * 3 + 255) / 256)), dim3(256), 0, s, d.as<double>(),
                       L.mask.as<uint8_t>(), L.dinv64.as<double>(), L.dinv32.as<float>(), nn);
    L.st32.ensure(sizeof(float) * 243 * (size_t)nn);
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((243 * nn + 255) / 256)), dim3(256), 0, s, L.st64.as<double>(),
                       L.st32.as<float>(), 243 * nn);
    prev = L.st64.as<double>();
    sync(h);
  }
  fst.release();
  ensure_workspace(h, 1);
  // side streams only for large grids: creating a stream costs milliseconds, more than the overlap gains on
  // small ones (many small handles, as in a batch of variants, keep everything on their own stream)
  const bool fork = F.g.nodes() >= (1LL << 20);
  if (fork)
    for (hipStream_t& q : h->side)
      if (!q) DFMI_HIP(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
  const hipStream_t side_inv = fork ? h->side[2] : s, side_l1 = fork ? h->side[0] : s, side_lc = fork ? h->side[1] : s;
  auto event_on = [](hipStream_t st) {
    hipEvent_t e;
    DFMI_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    DFMI_HIP(hipEventRecord(e, st));
    return e;
  };
  std::vector<hipEvent_t> joins;
  // coarsest dense inverse over live DOFs: assembled on the handle's stream, inverted on a side stream (clamped
  // fixtures) while the power iterations run
  Level& C = *H.levels.back();
  const long long nc = C.g.nodes();
  std::vector<uint8_t> mk(nc);
  DFMI_HIP(hipMemcpyAsync(mk.data(), C.mask.p, nc, hipMemcpyDeviceToHost, s));
  sync(h);
  std::vector<int> index(nc * 3, -1), live;
  for (long long i = 0; i < nc * 3; ++i)
    if ((mk[i / 3] >> (i % 3)) & 1) {
      index[i] = (int)live.size();
      live.push_back((int)i);
    }
  H.coarsest_n = (int)live.size();
  if (H.coarsest_n > DENSE_COARSEST_MAX)
    throw std::runtime_error("coarsest level too large for the dense solve; add levels");
  DevBuf idx;
  if (H.coarsest_n > 0) {
    idx.ensure(sizeof(int) * nc * 3);
    H.live_dofs.ensure(sizeof(int) * live.size());
    DFMI_HIP(hipMemcpyAsync(idx.p, index.data(), sizeof(int) * nc * 3, hipMemcpyHostToDevice, s));
    DFMI_HIP(hipMemcpyAsync(H.live_dofs.p, live.data(), sizeof(int) * live.size(), hipMemcpyHostToDevice, s));
    const int n = H.coarsest_n;
    H.ainv.ensure(sizeof(double) * (size_t)n * n);
    const double* cst = h->nlev > 1 ? C.st64.as<double>() : nullptr;
    if (!cst) throw std::runtime_error("multigrid needs at least two levels");
    coarsest_assemble(C.g, cst, idx.as<int>(), n, H.ainv.as<double>(), s);
    if (nk > 0) {
      // floating: pin nk DOFs, invert the rest, and fold the kernel projection into the inverse
      // (topology_multigrid.py:233-246, 267-283)
      const double scale = (double)(1 << (h->nlev - 1));
      const double hc[3] = {h->spacing[0] * scale, h->spacing[1] * scale, h->spacing[2] * scale};
      std::vector<double> kernel;
      std::vector<int> pins;
      coarsest_kernel(C.g, hc, center, coef, nk, live, kernel, pins);
      DevBuf pd;
      pd.ensure(sizeof(int) * pins.size());
      DFMI_HIP(hipMemcpyAsync(pd.p, pins.data(), sizeof(int) * pins.size(), hipMemcpyHostToDevice, s));
      pin_dense(H.ainv.as<double>(), n, pd.as<int>(), (int)pins.size(), s);
      dense_sweep_inverse(H.ainv.as<double>(), n, s);
      project_dense_inverse(H.ainv.as<double>(), n, pd.as<int>(), (int)pins.size(), kernel, nk, s);
      sync(h);
    } else {
      const hipEvent_t ready = event_on(s);
      DFMI_HIP(hipStreamWaitEvent(side_inv, ready, 0));
      DFMI_HIP(hipEventDestroy(ready));
      dense_sweep_inverse(H.ainv.as<double>(), n, side_inv);
      if (fork) H.ainv_ready = event_on(side_inv);  // joined by the first coarsest solve, not here
    }
  }
  // lambda_max of every level but the coarsest: level 0 on the handle's stream, level 1 and the coarser levels
  // on side streams; every buffer is allocated here, before the streams fork, and released after they join
  const int np = h->nlev - 1;
  std::vector<PowerBufs> pbs(np);
  for (int l = 0; l < np; ++l) {
    const long long n = col(H.levels[l]->g);
    pbs[l].v.ensure(sizeof(double) * n);
    pbs[l].y.ensure(sizeof(double) * n);
    pbs[l].nrm.ensure(sizeof(double) * 2);
    pbs[l].part.ensure(sizeof(double) * reduce_blocks(n));
  }
  {
    const hipEvent_t ready = event_on(s);
    DFMI_HIP(hipStreamWaitEvent(side_l1, ready, 0));
    DFMI_HIP(hipStreamWaitEvent(side_lc, ready, 0));
    DFMI_HIP(hipEventDestroy(ready));
  }
  for (int l = 1; l < np; ++l) spectrum_async(h, H, l, l == 1 ? side_l1 : side_lc, pbs[l]);
  if (np > 1) joins.push_back(event_on(side_l1));
  if (np > 2) joins.push_back(event_on(side_lc));
  if (np > 0) spectrum_async(h, H, 0, s, pbs[0]);
  for (hipEvent_t e : joins) {
    DFMI_HIP(hipStreamWaitEvent(s, e, 0));
    DFMI_HIP(hipEventDestroy(e));
  }
  std::vector<double> sqs(np, 0.0);
  for (int l = 0; l < np; ++l)
    DFMI_HIP(hipMemcpyAsync(&sqs[l], pbs[l].nrm.p, sizeof(double), hipMemcpyDeviceToHost, s));
  sync(h);
  for (int l = 0; l < np; ++l) H.levels[l]->lambda_max = std::sqrt(sqs[l]);
  if (nk > 0) kernel_check(h, H, coef, nk, center);
  H.built = true;
  H.setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---- PCG ----------------------------------------------------------------------------------
// x -= B (B^T x) for k fine-level columns, B the hierarchy's orthonormal rigid kernel
// (topology_multigrid.py:224-228)
static void remove_kernel(Handle* h, Hierarchy& H, double* x, int k) {
  if (H.nk == 0) return;
  const long long n = col(h->pad);
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int kc = std::min(64, k - c0);
    double* xc = x + (size_t)c0 * n;
    gram_device(H.kern.as<double>(), xc, n, H.nk, kc, H.kcoef.as<double>(), H.kwork.as<double>(), KERNEL_GRAM_BLOCKS,
                -1.0, h->stream);
    combine_device(H.kern.as<double>(), n, H.nk, H.kcoef.as<double>(), kc, xc, 1, h->stream);
  }
}

template <typename T>
static void precondition(Handle* h, Hierarchy& H, const double* r, double* z, int k, const int* active) {
  hipStream_t s = h->stream;
  Level& F = *H.levels[0];
  const long long n = col(F.g);
  ensure_level_vectors(h, H, k);
  maybe_tune_fine(h, H, k);
  VArgs v{h, &H, k, active};
  double* scale = h->scal.as<double>() + 8 * (size_t)h->k_cap;
  if (sizeof(T) == sizeof(double)) {
    vcycle<T>(v, 0, reinterpret_cast<const T*>(r));
    copy_columns<double>(reinterpret_cast<const double*>(F.x.as<T>()), z, n, k, active, s);
    return;
  }
  const int nb = reduce_blocks(n);
  column_absmax<double>(r, n, k, h->partial.as<double>(), nb, active, s);
  reduce_max(h->partial.as<double>(), nb, k, scale, s);
  to_work<T>(r, scale, F.rhs.as<T>(), n, k, active, s);
  vcycle<T>(v, 0, F.rhs.as<T>());
  from_work<T>(F.x.as<T>(), scale, z, n, k, active, s);
}

static void precond_any(Handle* h, Hierarchy& H, const double* r, double* z, int k, const int* active) {
  if (h->st.precision == 1)
    precondition<float>(h, H, r, z, k, active);
  else
    precondition<double>(h, H, r, z, k, active);
  remove_kernel(h, H, z, k);
}

static void fine_fp64(Handle* h, Hierarchy& H, int mode, const double* x, double* y, const double* b, int k,
                      const int* active, double* partial, int* nblocks) {
  ApplyArgs a{};
  a.x = x;
  a.y = y;
  a.b = b;
  a.k = k;
  a.active = active;
  a.partial = partial;
  a.premasked = 1;  // right-hand sides, start vectors and directions are masked before they reach the operator
  fine_apply<double>(h->pad, h->m64.as<double>(), H.levels[0]->mask.as<uint8_t>(), h->kb, mode, a, h->stream,
                     nblocks);
}

static void column_norms(Handle* h, const double* v, long long n, int k, const int* active, std::vector<double>& out) {
  const int nb = reduce_blocks(n);
  column_dot<double>(v, v, n, k, h->partial.as<double>(), nb, active, h->stream);
  reduce_partials(h->partial.as<double>(), nb, k, h->scal.as<double>(), h->stream);
  out.resize(k);
  DFMI_HIP(hipMemcpyAsync(out.data(), h->scal.as<double>(), sizeof(double) * k, hipMemcpyDeviceToHost, h->stream));
  sync(h);
  for (auto& x : out) x = std::sqrt(x);
}

// warm: start from the flat iterate in x_flat (masked) instead of zero
static void solve(Handle* h, int slot, const double* rhs_flat, double* x_flat, int k, dfmi_report* rep,
                  bool warm = false) {
  Hierarchy& H = h->slots[slot];
  if (!H.built) throw std::runtime_error("hierarchy slot not built");
  ensure_workspace(h, k);
  ensure_level_vectors(h, H, k);
  maybe_tune_fine(h, H, k);
  hipStream_t s = h->stream;
  const long long n = col(h->pad);
  double* b = h->b.as<double>();
  double* x = h->x.as<double>();
  double* r = h->r.as<double>();
  double* z = h->z.as<double>();
  double* zn = h->zn.as<double>();
  double* p = h->p.as<double>();
  double* ap = h->ap.as<double>();
  double* tmp = h->tmp.as<double>();
  double* sc = h->scal.as<double>();
  double* alpha = sc;
  double* rz = sc + (size_t)h->k_cap;
  double* pap = sc + 2 * (size_t)h->k_cap;
  double* rzn = sc + 3 * (size_t)h->k_cap;
  double* rzc = sc + 4 * (size_t)h->k_cap;
  double* beta = sc + 5 * (size_t)h->k_cap;
  double* rr = sc + 6 * (size_t)h->k_cap;
  int* act = h->active.as<int>();
  Level& F = *H.levels[0];

  flat_to_grid(h->pad, h->ux, h->uy, h->uz, rhs_flat, b, k, s);
  mask_vector(F.mask.as<uint8_t>(), b, h->pad.nodes(), k, s);
  remove_kernel(h, H, b, k);
  std::vector<double> nb;
  column_norms(h, b, n, k, nullptr, nb);
  std::vector<int> active(k);
  for (int c = 0; c < k; ++c) {
    active[c] = nb[c] > 0 ? 1 : 0;
    rep[c].iterations = 0;
    rep[c].relative_residual = 0.0;
    rep[c].converged = 1;
  }
  DFMI_HIP(hipMemcpyAsync(act, active.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
  if (warm) {
    flat_to_grid(h->pad, h->ux, h->uy, h->uz, x_flat, x, k, s);
    mask_vector(F.mask.as<uint8_t>(), x, h->pad.nodes(), k, s);
    fine_fp64(h, H, MODE_RESID, x, r, b, k, nullptr, nullptr, nullptr);
    remove_kernel(h, H, r, k);
  } else {
    fill_zero<double>(x, (long long)k * n, s);
    copy_columns<double>(b, r, n, k, nullptr, s);
  }
  precond_any(h, H, r, z, k, act);
  copy_columns<double>(z, p, n, k, nullptr, s);
  const int nbr = reduce_blocks(n);
  column_dot<double>(r, z, n, k, h->partial.as<double>(), nbr, act, s);
  reduce_partials(h->partial.as<double>(), nbr, k, rz, s);
  const double tol = h->st.tolerance;
  int step = 0;
  std::vector<double> rrh(k), rel(k);
  auto any = [&]() {
    for (int c = 0; c < k; ++c)
      if (active[c]) return true;
    return false;
  };
  while (any() && step < h->st.max_iterations) {
    int nbf = 0;
    fine_fp64(h, H, MODE_DOT, p, ap, nullptr, k, act, h->partial.as<double>(), &nbf);
    reduce_partials(h->partial.as<double>(), nbf, k, pap, s);
    pcg_scalars_alpha(rz, pap, alpha, k, act, s);
    pcg_update<double>(x, r, p, ap, alpha, n, k, act, h->partial.as<double>(), nbr, s);
    reduce_partials(h->partial.as<double>(), nbr, k, rr, s);
    DFMI_HIP(hipMemcpyAsync(rrh.data(), rr, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    sync(h);
    ++step;
    std::vector<int> check(k, 0);
    bool need = false;
    double worst = 0.0;
    for (int c = 0; c < k; ++c) {
      if (!active[c]) continue;
      rel[c] = std::sqrt(rrh[c]) / nb[c];
      worst = std::max(worst, rel[c]);
      if (rel[c] < tol) check[c] = 1, need = true;
    }
    h->history.push_back(worst);
    if (need) {
      // true residual confirmation (topology_multigrid.py:431-437)
      DFMI_HIP(hipMemcpyAsync(act, check.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
      fine_fp64(h, H, MODE_RESID, x, tmp, b, k, act, nullptr, nullptr);
      remove_kernel(h, H, tmp, k);
      std::vector<double> tn;
      column_norms(h, tmp, n, k, act, tn);
      std::vector<int> replace(k, 0);
      for (int c = 0; c < k; ++c) {
        if (!check[c]) continue;
        const double t = tn[c] / nb[c];
        if (t < tol) {
          active[c] = 0;
          rep[c].iterations = step;
          rep[c].relative_residual = t;
          rep[c].converged = 1;
        } else {
          replace[c] = 1;
        }
      }
      DFMI_HIP(hipMemcpyAsync(act, replace.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
      copy_columns<double>(tmp, r, n, k, act, s);
      DFMI_HIP(hipMemcpyAsync(act, active.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
    }
    if (!any()) break;
    precond_any(h, H, r, zn, k, act);
    column_dot<double>(r, zn, n, k, h->partial.as<double>(), nbr, act, s);
    reduce_partials(h->partial.as<double>(), nbr, k, rzn, s);
    column_dot<double>(r, z, n, k, h->partial.as<double>(), nbr, act, s);
    reduce_partials(h->partial.as<double>(), nbr, k, rzc, s);
    pcg_beta(rzn, rzc, rz, beta, k, act, s);
    pcg_direction<double>(p, zn, beta, n, k, act, s);
    std::swap(z, zn);
  }
  if (any()) {
    fine_fp64(h, H, MODE_RESID, x, tmp, b, k, act, nullptr, nullptr);
    remove_kernel(h, H, tmp, k);
    std::vector<double> tn;
    column_norms(h, tmp, n, k, act, tn);
    for (int c = 0; c < k; ++c)
      if (active[c]) {
        rep[c].iterations = step;
        rep[c].relative_residual = tn[c] / nb[c];
        rep[c].converged = rep[c].relative_residual < tol;
      }
  }
  // keep buffer roles stable for the next call
  if (z != h->z.as<double>()) std::swap(h->z.p, h->zn.p);
  grid_to_flat(h->pad, h->ux, h->uy, h->uz, x, x_flat, k, s);
  sync(h);
}

// Mixed-precision PCG with reliable updates.  The Krylov iteration runs entirely in FP32 (vectors,
// operator, V-cycle; dot products accumulated in FP64) on the scaled system A d = r / s.  A column makes a
// reliable update when its recursive FP32 residual has dropped by the factor `inner_tolerance` since
// the last update, or looks converged: the FP32 correction is flushed into the FP64 solution
// (x += s d, d = 0), the true residual r = b - A x is recomputed in FP64, and the iteration continues on
// r / |r| with its search direction, previous preconditioned residual and r.z rescaled to the new scale
// (no restart, so the Krylov information is kept).  Columns converge only on the FP64 true residual, so
// the result meets the same criterion as the flexible FP64 PCG above.
static void solve_refine(Handle* h, int slot, const double* rhs_flat, double* x_flat, int k, dfmi_report* rep) {
  if (h->st.precision != 1) throw std::invalid_argument("refine mode needs the fp32 V-cycle precision");
  Hierarchy& H = h->slots[slot];
  if (!H.built) throw std::runtime_error("hierarchy slot not built");
  ensure_workspace(h, k);
  ensure_level_vectors(h, H, k);
  maybe_tune_fine(h, H, k);
  hipStream_t s = h->stream;
  const long long n = col(h->pad);
  const size_t kn = (size_t)k * n;
  double* b = h->b.as<double>();
  double* x = h->x.as<double>();
  double* r = h->r.as<double>();
  // FP32 inner vectors live in the halves of FP64 buffers that this scheme does not use
  float* r32 = h->z.as<float>();
  float* d32 = r32 + kn;
  float* p32 = h->zn.as<float>();
  float* ap32 = p32 + kn;
  float* pn32 = reinterpret_cast<float*>(h->p.as<double>());  // next direction (formed inside the Krylov apply)
  double* sc = h->scal.as<double>();
  const size_t kc = (size_t)h->k_cap;
  double* alpha = sc;
  double* rz = sc + kc;
  double* pap = sc + 2 * kc;
  double* rzn = sc + 3 * kc;
  double* rzc = sc + 4 * kc;
  double* beta = sc + 5 * kc;
  double* rr = sc + 6 * kc;
  double* scale = sc + 8 * kc;
  double* factor = sc + 9 * kc;
  double* rzc_new = sc + 10 * kc;
  int* act = h->active.as<int>();
  Level& F = *H.levels[0];
  float* z = nullptr;  // current preconditioned residual: the fine level's V-cycle output
  VArgs v{h, &H, k, act};
  const int nbr = reduce_blocks(n);
  double* part = h->partial.as<double>();
  // r.z from the V-cycle's last fine smoothing step (vector kernel) instead of a separate pass
  const bool fused_rz = !use_mfma(h, F.g, MODE_CHEB);
  int nbc = 0;
  if (fused_rz) {
    if ((size_t)fine_grid_blocks<float>(F.g, MODE_CHEB, k) * k * sizeof(double) > h->partial.bytes)
      throw std::logic_error("partial-sum buffer too small for the fused r.z");
    v.final_dot = part;
    v.final_blocks = &nbc;
  }
  auto precondition_rz = [&](double* out) {  // z = V(r32), out = r32.z per column
    vcycle<float>(v, 0, r32);
    z = F.x.as<float>();
    if (fused_rz) {
      reduce_partials(part, nbc, k, out, s);
    } else {
      column_dot<float>(r32, z, n, k, part, nbr, act, s);
      reduce_partials(part, nbr, k, out, s);
    }
  };
  bool first = true;  // the first direction is z itself

  flat_to_grid(h->pad, h->ux, h->uy, h->uz, rhs_flat, b, k, s);
  mask_vector(F.mask.as<uint8_t>(), b, h->pad.nodes(), k, s);
  std::vector<double> nb;
  column_norms(h, b, n, k, nullptr, nb);
  std::vector<int> active(k), upd(k);
  std::vector<double> sh(k, 1.0), fh(k, 1.0), rrh(k), tn;
  for (int c = 0; c < k; ++c) {
    active[c] = nb[c] > 0 ? 1 : 0;
    sh[c] = nb[c] > 0 ? nb[c] : 1.0;
    rep[c].iterations = 0;
    rep[c].relative_residual = 0.0;
    rep[c].converged = 1;
  }
  const double tol = h->st.tolerance;
  const double delta = std::min(0.5, std::max(h->st.inner_tolerance, 1e-7));
  auto count = [](const std::vector<int>& f) {
    int m = 0;
    for (int q : f) m += q;
    return m;
  };
  auto upload = [&](int* dst, const std::vector<int>& src) {
    DFMI_HIP(hipMemcpyAsync(dst, src.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
  };
  fill_zero<double>(x, (long long)kn, s);
  fill_zero<float>(d32, (long long)kn, s);
  DFMI_HIP(hipMemcpyAsync(scale, sh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  upload(act, active);
  to_work<float>(b, scale, r32, n, k, act, s);
  int step = 0;
  if (count(active)) precondition_rz(rz);
  while (count(active) && step < h->st.max_iterations) {
    int nbf = 0;
    ApplyArgs a{};
    a.x = p32;  // previous direction (not read for the first one)
    a.b = z;
    a.beta = first ? nullptr : beta;
    a.pout = pn32;  // p = z + beta p_prev, formed while staging the operand
    a.y = ap32;
    a.k = k;
    a.active = act;
    a.partial = part;
    a.premasked = 1;  // z (V-cycle output) and the previous direction are masked
    fine_apply<float>(h->pad, h->m32.as<float>(), F.mask.as<uint8_t>(), h->kb, MODE_DOT, a, s, &nbf);
    std::swap(p32, pn32);
    first = false;
    reduce_partials(part, nbf, k, pap, s);
    pcg_scalars_alpha(rz, pap, alpha, k, act, s);
    // d += alpha p, r -= alpha Ap, |r|^2 and r_new . z_old in one pass
    pcg_update<float>(d32, r32, p32, ap32, alpha, n, k, act, part, nbr, s, z);
    reduce_partials(part, nbr, k, rr, s);
    reduce_partials(part + (size_t)k * nbr, nbr, k, rzc, s);
    DFMI_HIP(hipMemcpyAsync(rrh.data(), rr, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    sync(h);
    ++step;
    bool any_upd = false;
    double worst = 0.0;
    for (int c = 0; c < k; ++c) {
      upd[c] = 0;
      if (!active[c]) continue;
      const double est = std::sqrt(rrh[c]);  // |r| / s since the last update (r / s had norm 1 there)
      worst = std::max(worst, est * sh[c] / nb[c]);
      if (est * sh[c] < tol * nb[c] || est < delta) upd[c] = 1, any_upd = true;
    }
    h->history.push_back(worst);
    if (any_upd) {
      upload(act, upd);
      add_scaled(x, d32, scale, n, k, act, s, true);  // x += s d, then d = 0
      fine_fp64(h, H, MODE_RESID, x, r, b, k, act, nullptr, nullptr);
      column_norms(h, r, n, k, act, tn);
      std::vector<int> cont(k, 0);
      for (int c = 0; c < k; ++c) {
        fh[c] = 1.0;
        if (!upd[c]) continue;
        const double t = tn[c] / nb[c];
        rep[c].iterations = step;
        rep[c].relative_residual = t;
        rep[c].converged = t < tol;
        if (t < tol || !(tn[c] > 0)) {
          active[c] = 0;
        } else {
          cont[c] = 1;
          fh[c] = sh[c] / tn[c];
          sh[c] = tn[c];
        }
      }
      if (count(cont)) {
        upload(act, cont);
        DFMI_HIP(hipMemcpyAsync(scale, sh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        DFMI_HIP(hipMemcpyAsync(factor, fh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        to_work<float>(r, scale, r32, n, k, act, s);  // r / |r|
        scale_columns<float>(p32, n, k, factor, act, s);
        scale_columns<float>(z, n, k, factor, act, s);
        select_scalars(factor, rz, act, k, true, s);  // rz *= f^2
        column_dot<float>(r32, z, n, k, part, nbr, act, s);
        reduce_partials(part, nbr, k, rzc_new, s);
        select_scalars(rzc_new, rzc, act, k, false, s);
      }
      upload(act, active);
    }
    if (!count(active)) break;
    precondition_rz(rzn);
    pcg_beta(rzn, rzc, rz, beta, k, act, s);  // the direction itself is formed by the next Krylov apply
  }
  if (count(active)) {  // iteration limit: flush and report the true residual
    upload(act, active);
    add_scaled(x, d32, scale, n, k, act, s);
    fine_fp64(h, H, MODE_RESID, x, r, b, k, act, nullptr, nullptr);
    column_norms(h, r, n, k, act, tn);
    for (int c = 0; c < k; ++c)
      if (active[c]) {
        rep[c].iterations = step;
        rep[c].relative_residual = tn[c] / nb[c];
        rep[c].converged = rep[c].relative_residual < tol;
      }
  }
  grid_to_flat(h->pad, h->ux, h->uy, h->uz, x, x_flat, k, s);
  sync(h);
}

// ---- block Krylov solve -------------------------------------------------------------------
constexpr int BLOCK_MAX_K = 48;

// Pseudo-inverse of the block Gram matrix G = V^T A V (k x k, row-major, symmetrised here) over a
// well-conditioned subset of the columns: with C = D^-1/2 G D^-1/2 (unit diagonal), a Cholesky factorisation
// with symmetric pivoting on the largest remaining diagonal keeps columns while their A-energy orthogonal
// to the kept ones, relative to their own, exceeds `drop`; then S = D^-1/2 Pi (L L^T)^-1 Pi^T D^-1/2 on the
// kept columns and 0 elsewhere.  S = T T^T with T = D^-1/2 Pi L^-T, and V T is A-orthonormal.  Returns the
// number of kept columns.  (k <= 48: a few tens of microseconds on the host.)
static int block_pinv(const std::vector<double>& g, int k, double drop, std::vector<double>& S) {
  S.assign((size_t)k * k, 0.0);
  std::vector<int> idx;
  for (int c = 0; c < k; ++c) {
    const double d = g[(size_t)c * k + c];
    if (d > 0 && std::isfinite(d)) idx.push_back(c);
  }
  const int nz = (int)idx.size();
  if (nz == 0) return 0;
  std::vector<double> dn(nz), C((size_t)nz * nz);
  for (int i = 0; i < nz; ++i) dn[i] = std::sqrt(g[(size_t)idx[i] * k + idx[i]]);
  for (int i = 0; i < nz; ++i)
    for (int j = 0; j < nz; ++j)
      C[(size_t)i * nz + j] =
          0.5 * (g[(size_t)idx[i] * k + idx[j]] + g[(size_t)idx[j] * k + idx[i]]) / (dn[i] * dn[j]);
  std::vector<int> perm(nz);
  for (int i = 0; i < nz; ++i) perm[i] = i;
  auto at = [&](int i, int j) -> double& { return C[(size_t)i * nz + j]; };
  int m = 0;
  for (; m < nz; ++m) {
    int piv = m;
    for (int q = m + 1; q < nz; ++q)
      if (at(q, q) > at(piv, piv)) piv = q;
    if (!(at(piv, piv) > drop)) break;
    if (piv != m) {  // symmetric swap of rows and columns m, piv
      for (int q = 0; q < nz; ++q) std::swap(at(m, q), at(piv, q));
      for (int q = 0; q < nz; ++q) std::swap(at(q, m), at(q, piv));
      std::swap(perm[m], perm[piv]);
    }
    const double lmm = std::sqrt(at(m, m));
    at(m, m) = lmm;
    for (int i = m + 1; i < nz; ++i) at(i, m) /= lmm;
    for (int i = m + 1; i < nz; ++i)
      for (int j = m + 1; j <= i; ++j) {
        at(i, j) -= at(i, m) * at(j, m);
        at(j, i) = at(i, j);
      }
  }
  if (m == 0) return 0;
  // Li = L^-1 (m x m lower triangular) by forward substitution, then (L L^T)^-1 = Li^T Li
  std::vector<double> Li((size_t)m * m, 0.0);
  for (int c = 0; c < m; ++c) {
    Li[(size_t)c * m + c] = 1.0 / at(c, c);
    for (int i = c + 1; i < m; ++i) {
      double acc = 0.0;
      for (int q = c; q < i; ++q) acc += at(i, q) * Li[(size_t)q * m + c];
      Li[(size_t)i * m + c] = -acc / at(i, i);
    }
  }
  for (int a = 0; a < m; ++a)
    for (int b = 0; b <= a; ++b) {
      double acc = 0.0;
      for (int q = a; q < m; ++q) acc += Li[(size_t)q * m + a] * Li[(size_t)q * m + b];
      const int ia = perm[a], ib = perm[b];
      const double v = acc / (dn[ia] * dn[ib]);
      S[(size_t)idx[ia] * k + idx[ib]] = v;
      S[(size_t)idx[ib] * k + idx[ia]] = v;
    }
  return m;
}

// Breakdown-free block PCG over the k columns (FP32 vectors and V-cycle, reliable FP64 updates as in
// solve_refine).  With W = A V, G = V^T W and S its pseudo-inverse (block_pinv), the A-orthonormal block
// P = V T (S = T T^T) is never formed: the step is D += V a, R -= W a with a = S V^T R, and the next block is
// V' = Z - V S W^T Z for Z = M R (= Z + P beta, beta = -(A P)^T Z).  One iteration therefore makes four passes
// over the DOFs besides the operator apply and the V-cycle: the Gram V^T [W | R], the paired update (which
// also returns |R_j|^2), the Gram W^T Z and the combination (block.hip).  Only S (k x k) is formed on the host.
static void solve_block(Handle* h, int slot, const double* rhs_flat, double* x_flat, int k, dfmi_report* rep) {
  if (h->st.precision != 1) throw std::invalid_argument("block Krylov needs the fp32 V-cycle precision");
  if (k > BLOCK_MAX_K) throw std::invalid_argument("block Krylov takes at most 48 columns");
  Hierarchy& H = h->slots[slot];
  if (!H.built) throw std::runtime_error("hierarchy slot not built");
  ensure_workspace(h, k);
  ensure_level_vectors(h, H, k);
  hipStream_t s = h->stream;
  const long long n = col(h->pad);
  const size_t kn = (size_t)k * n;
  double* b = h->b.as<double>();
  double* x = h->x.as<double>();
  double* r = h->r.as<double>();
  float* r32 = h->z.as<float>();
  float* d32 = r32 + kn;
  float* vcur = h->zn.as<float>();  // V and the next V alternate between the halves of zn
  float* vnext = vcur + kn;
  float* w32 = reinterpret_cast<float*>(h->p.as<double>());
  const int gblocks = block_gram_blocks(n), cblocks = block_combine_blocks(n);
  h->bk_work.ensure(sizeof(double) * std::max((size_t)BLOCK_MAX_K * 2 * BLOCK_MAX_K * gblocks,
                                              (size_t)BLOCK_MAX_K * cblocks));
  h->bk_mat.ensure(sizeof(double) * BLOCK_MAX_K * BLOCK_MAX_K * 6);
  double* g1 = h->bk_mat.as<double>();                  // k x 2k: V^T [W | R]
  double* smat = g1 + 2 * BLOCK_MAX_K * BLOCK_MAX_K;     // k x k: S
  double* hmat = smat + BLOCK_MAX_K * BLOCK_MAX_K;       // k x k: W^T Z
  float* amat = reinterpret_cast<float*>(hmat + BLOCK_MAX_K * BLOCK_MAX_K);  // k x 48 FP32: step coefficients
  float* bmat = amat + BLOCK_MAX_K * BLOCK_MAX_K;                            // k x 48 FP32: next-block coefficients
  double* work = h->bk_work.as<double>();
  double* sc = h->scal.as<double>();
  const size_t kc = (size_t)h->k_cap;
  double* rr = sc + 6 * kc;
  double* scale = sc + 8 * kc;
  double* factor = sc + 9 * kc;
  double* invs = sc + 10 * kc;  // per-column factor of a restarted FP32 residual
  double* tnd = sc + 11 * kc;   // |r|^2 of the flushed columns
  int* act = h->active.as<int>();
  Level& F = *H.levels[0];
  VArgs v{h, &H, k, act};

  flat_to_grid(h->pad, h->ux, h->uy, h->uz, rhs_flat, b, k, s);
  mask_vector(F.mask.as<uint8_t>(), b, h->pad.nodes(), k, s);
  std::vector<double> nb;
  column_norms(h, b, n, k, nullptr, nb);
  std::vector<int> active(k), upd(k), inactive(k);
  std::vector<double> sh(k, 1.0), fh(k, 0.0), rrh(k), tn;
  for (int c = 0; c < k; ++c) {
    active[c] = nb[c] > 0 ? 1 : 0;
    sh[c] = nb[c] > 0 ? nb[c] : 1.0;
    rep[c].iterations = 0;
    rep[c].relative_residual = 0.0;
    rep[c].converged = 1;
  }
  const double tol = h->st.tolerance;
  const double delta = std::min(0.5, std::max(h->st.inner_tolerance, 1e-7));
  auto count = [](const std::vector<int>& f) {
    int m = 0;
    for (int q : f) m += q;
    return m;
  };
  auto upload = [&](int* dst, const std::vector<int>& src) {
    DFMI_HIP(hipMemcpyAsync(dst, src.data(), sizeof(int) * k, hipMemcpyHostToDevice, s));
  };
  // zero the columns whose flag in `which` is set (a write-only fill of the selected columns)
  auto zero_columns = [&](float* vec, const std::vector<int>& which) {
    if (!count(which)) return;
    upload(act, which);
    scale_columns<float>(vec, n, k, nullptr, act, s);
    upload(act, active);
  };
  auto inactive_of = [&]() {
    for (int c = 0; c < k; ++c) inactive[c] = active[c] ? 0 : 1;
    return inactive;
  };
  fill_zero<double>(x, (long long)kn, s);
  fill_zero<float>(d32, (long long)kn, s);
  fill_zero<float>(r32, (long long)kn, s);
  DFMI_HIP(hipMemcpyAsync(scale, sh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  upload(act, active);
  to_work<float>(b, scale, r32, n, k, act, s);
  if (!count(active)) {
    grid_to_flat(h->pad, h->ux, h->uy, h->uz, x, x_flat, k, s);
    sync(h);
    return;
  }
  // Z = M R, with the columns of finished right-hand sides zero (in the fine level's x vector)
  auto precondition_block = [&]() -> float* {
    vcycle<float>(v, 0, r32);
    float* z = F.x.as<float>();
    zero_columns(z, inactive_of());
    return z;
  };
  {
    const float* z = precondition_block();
    DFMI_HIP(hipMemcpyAsync(vcur, z, sizeof(float) * kn, hipMemcpyDeviceToDevice, s));
  }
  int step = 0;
  std::vector<double> g((size_t)k * 2 * k), S;
  std::vector<float> af((size_t)k * BLOCK_MAX_K);
  while (count(active) && step < h->st.max_iterations) {
    {  // W = A V over all k columns
      ApplyArgs a{};
      a.x = vcur;
      a.y = w32;
      a.k = k;
      a.premasked = 1;  // V is a combination of masked V-cycle outputs
      fine_apply<float>(h->pad, h->m32.as<float>(), F.mask.as<uint8_t>(), h->kb, MODE_APPLY, a, s, nullptr);
    }
    block_gram(vcur, k, w32, k, r32, k, n, g1, work, s);  // [V^T W | V^T R]
    DFMI_HIP(hipMemcpyAsync(g.data(), g1, sizeof(double) * k * 2 * k, hipMemcpyDeviceToHost, s));
    sync(h);
    std::vector<double> gw((size_t)k * k);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) gw[(size_t)i * k + j] = g[(size_t)i * 2 * k + j];
    if (block_pinv(gw, k, 1e-5, S) == 0) break;
    // a = S V^T R (rows: input columns of V, columns: right-hand sides)
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < BLOCK_MAX_K; ++j) {
        double acc = 0.0;
        if (j < k)
          for (int q = 0; q < k; ++q) acc += S[(size_t)i * k + q] * g[(size_t)q * 2 * k + k + j];
        af[(size_t)i * BLOCK_MAX_K + j] = (float)acc;
      }
    DFMI_HIP(hipMemcpyAsync(amat, af.data(), sizeof(float) * k * BLOCK_MAX_K, hipMemcpyHostToDevice, s));
    DFMI_HIP(hipMemcpyAsync(smat, S.data(), sizeof(double) * k * k, hipMemcpyHostToDevice, s));
    // R -= W a, |R_j|^2; D += V a waits for the pass that forms the next block (it reads V as well), unless a
    // column is flushed below
    block_residual_update(w32, r32, amat, k, n, rr, work, s);
    DFMI_HIP(hipMemcpyAsync(rrh.data(), rr, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    sync(h);
    ++step;
    bool any_upd = false;
    double worst = 0.0;
    for (int c = 0; c < k; ++c) {
      upd[c] = 0;
      if (!active[c]) continue;
      const double est = std::sqrt(rrh[c]);
      worst = std::max(worst, est * sh[c] / nb[c]);
      if (est * sh[c] < tol * nb[c] || est < delta) upd[c] = 1, any_upd = true;
    }
    h->history.push_back(worst);
    // a flush also takes every column whose residual has dropped by sqrt(inner_tolerance) since its last
    // restart: fewer flush iterations (each costs a separate pass for D += V a) and fewer restarts of the block
    if (any_upd)
      for (int c = 0; c < k; ++c)
        if (active[c] && !upd[c] && std::sqrt(rrh[c]) < std::sqrt(delta)) upd[c] = 1;
    bool pending = true;  // D += V a not yet applied
    if (any_upd) {  // flush D into x, true FP64 residual, restart the column's scaled residual at norm 1
      pending = false;
      upload(act, upd);
      // D += V a over all columns; the flushed ones move into x (x += s D) and restart D at zero in the same pass
      block_combine_flush(vcur, d32, amat, k, n, x, scale, act, s);
      // true FP64 residual of the flushed columns, written straight into R in FP32 scaled by the recursive estimate
      // of its norm, with the partial sums of |r|^2; the continuing columns are then brought to unit norm
      std::vector<double> shn(sh), inv(k, 1.0);
      for (int c = 0; c < k; ++c) {
        if (!upd[c]) continue;
        const double e = std::sqrt(rrh[c]) * sh[c];
        shn[c] = e > 1e-300 ? e : 1.0;
        inv[c] = 1.0 / shn[c];
      }
      DFMI_HIP(hipMemcpyAsync(invs, inv.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
      {
        ApplyArgs a{};
        a.x = x;
        a.b = b;
        a.k = k;
        a.active = act;
        a.partial = h->partial.as<double>();
        a.beta = invs;
        a.pout = r32;
        a.premasked = 1;
        int nbk = 0;
        fine_apply<double>(h->pad, h->m64.as<double>(), F.mask.as<uint8_t>(), h->kb, MODE_RESID_WORK, a, s, &nbk);
        reduce_partials(h->partial.as<double>(), nbk, k, tnd, s);
        tn.resize(k);
        DFMI_HIP(hipMemcpyAsync(tn.data(), tnd, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        sync(h);
        for (auto& t : tn) t = std::sqrt(t);
      }
      std::vector<int> cont(k, 0), done(k, 0);
      for (int c = 0; c < k; ++c) {
        fh[c] = 1.0;
        if (!upd[c]) continue;
        const double tr = tn[c] / nb[c];
        rep[c].iterations = step;
        rep[c].relative_residual = tr;
        rep[c].converged = tr < tol;
        if (tr < tol || !(tn[c] > 0)) {
          active[c] = 0;
          done[c] = 1;
        } else {  // restart at unit norm: r / |r|
          cont[c] = 1;
          fh[c] = shn[c] / tn[c];
          sh[c] = tn[c];
        }
      }
      if (count(cont)) {
        DFMI_HIP(hipMemcpyAsync(factor, fh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        upload(act, cont);
        scale_columns<float>(r32, n, k, factor, act, s);
      }
      DFMI_HIP(hipMemcpyAsync(scale, sh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
      zero_columns(r32, done);
      upload(act, active);
    }
    if (!count(active)) break;
    // Z = M R; V' = Z - V S W^T Z, with the deferred D += V a in the same pass over V
    const float* z = precondition_block();
    block_gram(w32, k, z, k, nullptr, 0, n, hmat, work, s);
    block_small_product(smat, hmat, k, -1.0, bmat, s);
    if (pending)
      block_combine2(vcur, z, vnext, bmat, d32, d32, amat, k, n, s);
    else
      block_combine(vcur, z, vnext, bmat, k, k, n, s);
    std::swap(vcur, vnext);
  }
  if (count(active)) {  // iteration limit: flush and report the true residual
    upload(act, active);
    add_scaled(x, d32, scale, n, k, act, s);
    fine_fp64(h, H, MODE_RESID, x, r, b, k, act, nullptr, nullptr);
    column_norms(h, r, n, k, act, tn);
    for (int c = 0; c < k; ++c)
      if (active[c]) {
        rep[c].iterations = step;
        rep[c].relative_residual = tn[c] / nb[c];
        rep[c].converged = rep[c].relative_residual < tol;
      }
  }
  grid_to_flat(h->pad, h->ux, h->uy, h->uz, x, x_flat, k, s);
  sync(h);
}

}  // namespace dfmi

// ---- C ABI ---------------------------------------------------------------------------------
using namespace dfmi;

struct dfmi_mg {
  Handle h;
};

struct dfmi_net {
  NetImpl* impl = nullptr;
};

#define DFMI_GUARD(...)                   \
  try {                                   \
    __VA_ARGS__;                          \
    return 0;                             \
  } catch (const std::exception& e) {     \
    dfmi::set_error(e.what());            \
    return 1;                             \
  } catch (...) {                         \
    dfmi::set_error("unknown failure");   \
    return 1;                             \
  }

// entry points of a handle: device buffers are allocated and released on the handle's stream
#define DFMI_GUARD_MG(...)                                     \
  try {                                                        \
    if (!mg) throw std::invalid_argument("null handle");       \
    dfmi::StreamScope dfmi_scope_(mg->h.stream);               \
    __VA_ARGS__;                                               \
    return 0;                                                  \
  } catch (const std::exception& e) {                          \
    dfmi::set_error(e.what());                                 \
    return 1;                                                  \
  } catch (...) {                                              \
    dfmi::set_error("unknown failure");                        \
    return 1;                                                  \
  }

extern "C" {

const char* dfmi_last_error(void) { return dfmi::get_error(); }

void dfmi_default_settings(dfmi_settings* o) {
  std::memset(o, 0, sizeof(*o));
  o->coarsest_dofs = 80000;
  o->max_levels = 8;
  o->levels = 0;
  o->chebyshev_degree = 3;
  o->chebyshev_ratio = 20.0;
  o->growth = 2.0;
  o->power_iterations = 20;
  o->precision = 1;
  o->tolerance = 1e-8;
  o->max_iterations = 2000;
  o->dead_ratio = 1e-12;
  o->seed = 0;
  o->refine = 0;
  o->inner_tolerance = 1e-2;
  o->fine_smoother = 0;
  o->retries = 2;
  o->coarse_krylov = 0;
  o->block_krylov = 0;
  o->smoother = 0;
  o->sweeps = 2;
  o->damping = 1.0;
  o->cycle = 0;
  o->coarse = 0;
  o->krylov_levels = 1;
  o->chebyshev_kind = 1;
}

int dfmi_create(const int32_t shape[3], const double spacing[3], const double* ke, const dfmi_settings* settings,
                dfmi_mg** out) {
  DFMI_GUARD({
    if (!shape || !spacing || !ke || !out) throw std::invalid_argument("null argument");
    keep_pool_memory();
    StreamScope scope(nullptr);  // the handle's buffers come from the stream-ordered pool as well
    for (int a = 0; a < 3; ++a) {
      if (shape[a] < 1) throw std::invalid_argument("grid shape must contain three positive integers");
      if (!(spacing[a] > 0) || !std::isfinite(spacing[a])) throw std::invalid_argument("spacing must be positive");
    }
    auto mg = std::make_unique<dfmi_mg>();
    Handle& h = mg->h;
    dfmi_settings st;
    dfmi_default_settings(&st);
    if (settings) st = *settings;
    if (st.precision != 0 && st.precision != 1) throw std::invalid_argument("precision must be 0 or 1");
    if (st.fine_smoother != 0 && st.fine_smoother != 1) throw std::invalid_argument("fine_smoother must be 0 or 1");
    if (st.retries < 0) throw std::invalid_argument("retries must be >= 0");
    if (st.coarse_krylov < 0 || st.coarse_krylov > 2) throw std::invalid_argument("coarse_krylov must be 0, 1 or 2");
    if (st.krylov_levels < 1) throw std::invalid_argument("krylov_levels must be >= 1");
    if (st.chebyshev_kind != 1 && st.chebyshev_kind != 4) throw std::invalid_argument("chebyshev_kind must be 1 or 4");
    if (st.block_krylov != 0 && st.block_krylov != 1) throw std::invalid_argument("block_krylov must be 0 or 1");
    if (st.chebyshev_degree < 1 || st.max_iterations < 1 || !(st.tolerance > 0))
      throw std::invalid_argument("invalid multigrid settings");
    if (st.smoother != 0 && st.smoother != 1) throw std::invalid_argument("smoother must be 0 (chebyshev) or 1 (jacobi)");
    if (st.smoother == 1 && (st.sweeps < 1 || !(st.damping > 0)))
      throw std::invalid_argument("jacobi needs sweeps >= 1 and damping > 0");
    if (st.cycle != 0 && st.cycle != 1) throw std::invalid_argument("cycle must be 0 (V) or 1 (W)");
    if (st.coarse != 0 && st.coarse != 1) throw std::invalid_argument("coarse must be 0 (galerkin) or 1 (rediscretize)");
    h.st = st;
    h.ux = shape[0];
    h.uy = shape[1];
    h.uz = shape[2];
    int sh[3] = {shape[0], shape[1], shape[2]}, padded[3];
    h.nlev = level_count(sh, st, padded);
    if (h.nlev < 2) throw std::invalid_argument("multigrid needs at least two levels");
    h.pad = make_geom(padded[0], padded[1], padded[2]);
    for (int l = 0; l < h.nlev; ++l)
      h.geoms.push_back(make_geom(padded[0] >> l, padded[1] >> l, padded[2] >> l));
    for (int a = 0; a < 3; ++a) h.spacing[a] = spacing[a];
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        const int qi = kRefToBits[i / 3] * 3 + i % 3, qj = kRefToBits[j / 3] * 3 + j % 3;
        h.ke_bits[qi * 24 + qj] = ke[i * 24 + j];
      }
    hex8_blocks(h.ke_bits, h.kb);
    double kd[24];
    for (int i = 0; i < 24; ++i) kd[i] = h.ke_bits[i * 24 + i];
    h.ke_dev.ensure(sizeof(double) * 576);
    h.kediag_dev.ensure(sizeof(double) * 24);
    DFMI_HIP(hipMemcpy(h.ke_dev.p, h.ke_bits, sizeof(double) * 576, hipMemcpyHostToDevice));
    DFMI_HIP(hipMemcpy(h.kediag_dev.p, kd, sizeof(double) * 24, hipMemcpyHostToDevice));
    std::vector<unsigned> tab;
    fine_mfma_table(h.ke_bits, tab);
    h.atab.ensure(sizeof(unsigned) * tab.size());
    DFMI_HIP(hipMemcpy(h.atab.p, tab.data(), sizeof(unsigned) * tab.size(), hipMemcpyHostToDevice));
    h.m64.ensure(sizeof(double) * h.pad.elems());
    h.m32.ensure(sizeof(float) * h.pad.elems());
    *out = mg.release();
  })
}

int dfmi_destroy(dfmi_mg* mg) {
  DFMI_GUARD({
    if (mg) {
      StreamScope scope(mg->h.stream);
      (void)hipStreamSynchronize(mg->h.stream);
      delete mg;
    }
  })
}

int dfmi_set_stream(dfmi_mg* mg, void* stream) { DFMI_GUARD_MG({ mg->h.stream = (hipStream_t)stream; }) }

int dfmi_levels(dfmi_mg* mg, int32_t* levels, int32_t padded[3]) {
  DFMI_GUARD_MG({
    *levels = mg->h.nlev;
    padded[0] = mg->h.pad.ex;
    padded[1] = mg->h.pad.ey;
    padded[2] = mg->h.pad.ez;
  })
}

int dfmi_update(dfmi_mg* mg, const double* moduli) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    pad_moduli(h.pad, h.ux, h.uy, h.uz, moduli, h.m64.as<double>(), h.m32.as<float>(), h.stream);
    for (auto& H : h.slots) retire(H, h.stream), H = Hierarchy();
    h.have_moduli = true;
    sync(&h);
  })
}

int dfmi_build_hierarchy(dfmi_mg* mg, int32_t slot, const uint8_t* fixed, double* setup_seconds) {
  DFMI_GUARD_MG({
    if (slot < 0 || slot >= 8) throw std::invalid_argument("slot must be in [0, 8)");
    build(&mg->h, slot, fixed);
    if (setup_seconds) *setup_seconds = mg->h.slots[slot].setup_s;
  })
}

int dfmi_build_hierarchy_kernel(dfmi_mg* mg, int32_t slot, const uint8_t* fixed, const double* coefficients,
                                int32_t nk, double* setup_seconds) {
  DFMI_GUARD_MG({
    if (slot < 0 || slot >= 8) throw std::invalid_argument("slot must be in [0, 8)");
    if (nk < 0 || nk > 6 || (nk > 0 && !coefficients)) throw std::invalid_argument("nk must be in [0, 6]");
    build(&mg->h, slot, fixed, coefficients, nk);
    if (setup_seconds) *setup_seconds = mg->h.slots[slot].setup_s;
  })
}

int dfmi_kernel_component(dfmi_mg* mg, int32_t slot, const double* rhs, int32_t k, int32_t* nk, double* component) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (slot < 0 || slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
    Hierarchy& H = h.slots[slot];
    if (nk) *nk = H.nk;
    if (!rhs || k < 1 || !component) return 0;
    ensure_workspace(&h, k);
    const long long n = col(h.pad);
    flat_to_grid(h.pad, h.ux, h.uy, h.uz, rhs, h.b.as<double>(), k, h.stream);
    mask_vector(H.levels[0]->mask.as<uint8_t>(), h.b.as<double>(), h.pad.nodes(), k, h.stream);
    std::vector<double> nb;
    column_norms(&h, h.b.as<double>(), n, k, nullptr, nb);
    for (int c = 0; c < k; ++c) component[c] = 0.0;
    if (H.nk == 0) return 0;
    for (int c0 = 0; c0 < k; c0 += 64) {
      const int kc = std::min(64, k - c0);
      gram_device(H.kern.as<double>(), h.b.as<double>() + (size_t)c0 * n, n, H.nk, kc, H.kcoef.as<double>(),
                  H.kwork.as<double>(), KERNEL_GRAM_BLOCKS, 1.0, h.stream);
      std::vector<double> cf((size_t)H.nk * kc);
      DFMI_HIP(hipMemcpyAsync(cf.data(), H.kcoef.p, sizeof(double) * cf.size(), hipMemcpyDeviceToHost, h.stream));
      sync(&h);
      for (int c = 0; c < kc; ++c) {
        double sq = 0.0;
        for (int a = 0; a < H.nk; ++a) sq += cf[(size_t)a * kc + c] * cf[(size_t)a * kc + c];
        component[c0 + c] = std::sqrt(sq) / std::max(nb[c0 + c], 1e-300);
      }
    }
  })
}

// Retry of failed columns in FP64 (topology_multigrid.py:376-406, _retried / _precise): up to `retries`
// times, the columns whose relative residual is not below the tolerance are solved again by flexible
// PCG with the FP64 V-cycle, from their previous iterate when its relative residual is below 1 and from
// zero otherwise; iterations add up and the relative residual is replaced.
static void retry_failed(Handle* h, int slot, const double* rhs, double* x, int k, dfmi_report* rep) {
  const long long nflat = 3LL * (h->ux + 1) * (h->uy + 1) * (h->uz + 1);
  for (int c = 0; c < k; ++c) rep[c].retried = 0;
  for (int round = 0; round < h->st.retries; ++round) {
    std::vector<int> failed;
    for (int c = 0; c < k; ++c)
      if (!(rep[c].relative_residual < h->st.tolerance)) failed.push_back(c);
    if (failed.empty()) return;
    const int nf = (int)failed.size();
    DevBuf rb, xb;
    rb.ensure(sizeof(double) * nf * nflat);
    xb.ensure(sizeof(double) * nf * nflat);
    const size_t colb = sizeof(double) * (size_t)nflat;
    for (int j = 0; j < nf; ++j) {
      const int c = failed[j];
      DFMI_HIP(hipMemcpyAsync(rb.as<double>() + (size_t)j * nflat, rhs + (size_t)c * nflat, colb,
                              hipMemcpyDeviceToDevice, h->stream));
      if (rep[c].relative_residual < 1.0)
        DFMI_HIP(hipMemcpyAsync(xb.as<double>() + (size_t)j * nflat, x + (size_t)c * nflat, colb,
                                hipMemcpyDeviceToDevice, h->stream));
      else
        fill_zero<double>(xb.as<double>() + (size_t)j * nflat, nflat, h->stream);
    }
    std::vector<dfmi_report> again(nf);
    const dfmi_settings saved = h->st;
    h->st.precision = 0;
    h->st.refine = 0;
    try {
      solve(h, slot, rb.as<double>(), xb.as<double>(), nf, again.data(), true);
    } catch (...) {
      h->st = saved;
      throw;
    }
    h->st = saved;
    for (int j = 0; j < nf; ++j) {
      const int c = failed[j];
      DFMI_HIP(hipMemcpyAsync(x + (size_t)c * nflat, xb.as<double>() + (size_t)j * nflat, colb,
                              hipMemcpyDeviceToDevice, h->stream));
      rep[c].iterations += again[j].iterations;
      rep[c].relative_residual = again[j].relative_residual;
      rep[c].converged = again[j].relative_residual < h->st.tolerance;
      rep[c].retried += 1;
    }
    sync(h);
  }
}

int dfmi_solve(dfmi_mg* mg, int32_t slot, const double* rhs, double* x, int32_t k, dfmi_report* report) {
  DFMI_GUARD_MG({
    if (slot < 0 || slot >= 8) throw std::invalid_argument("slot must be in [0, 8)");
    if (k < 1 || !rhs || !x || !report) throw std::invalid_argument("solve needs k >= 1 and buffers");
    mg->h.history.clear();
    if (mg->h.st.refine && mg->h.slots[slot].nk == 0 && mg->h.st.block_krylov && k <= BLOCK_MAX_K)
      solve_block(&mg->h, slot, rhs, x, k, report);
    else if (mg->h.st.refine && mg->h.slots[slot].nk == 0)
      solve_refine(&mg->h, slot, rhs, x, k, report);
    else
      solve(&mg->h, slot, rhs, x, k, report);
    retry_failed(&mg->h, slot, rhs, x, k, report);
  })
}

int dfmi_apply(dfmi_mg* mg, int32_t slot, int32_t precision, const double* x, double* y, int32_t k) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (!h.have_moduli) throw std::runtime_error("update(moduli) must precede apply");
    if (k < 1) throw std::invalid_argument("k must be >= 1");
    ensure_workspace(&h, k);
    const long long n = col(h.pad);
    hipStream_t s = h.stream;
    DevBuf unmasked;
    const uint8_t* mask;
    if (slot >= 0) {
      if (slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
      mask = h.slots[slot].levels[0]->mask.as<uint8_t>();
    } else {
      unmasked.ensure(h.pad.nodes());
      active_mask(h.pad, h.m64.as<double>(), nullptr, unmasked.as<uint8_t>(), s);
      mask = unmasked.as<uint8_t>();
    }
    flat_to_grid(h.pad, h.ux, h.uy, h.uz, x, h.b.as<double>(), k, s);
    ApplyArgs a{};
    a.k = k;
    if (precision == 0) {
      a.x = h.b.as<double>();
      a.y = h.tmp.as<double>();
      fine_apply<double>(h.pad, h.m64.as<double>(), mask, h.kb, MODE_APPLY, a, s, nullptr);
    } else {
      DevBuf xf, yf;
      xf.ensure(sizeof(float) * k * n);
      yf.ensure(sizeof(float) * k * n);
      hipLaunchKernelGGL(k_to_float, dim3((unsigned)((k * n + 255) / 256)), dim3(256), 0, s, h.b.as<double>(),
                         xf.as<float>(), k * n);
      a.x = xf.as<float>();
      a.y = yf.as<float>();
      if (precision == 2) {
        if (!fine_mfma_supported(h.pad, MODE_APPLY)) throw std::invalid_argument("grid not supported on the matrix cores");
        fine_apply_mfma(h.pad, h.m32.as<float>(), mask, h.atab.as<unsigned>(), MODE_APPLY, a, fine_tile(&h, k), s);
      } else {
        fine_apply<float>(h.pad, h.m32.as<float>(), mask, h.kb, MODE_APPLY, a, s, nullptr);
      }
      std::vector<double> ones(k, 1.0);
      DFMI_HIP(hipMemcpyAsync(h.scal.as<double>(), ones.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
      from_work<float>(yf.as<float>(), h.scal.as<double>(), h.tmp.as<double>(), n, k, nullptr, s);
      sync(&h);
    }
    grid_to_flat(h.pad, h.ux, h.uy, h.uz, h.tmp.as<double>(), y, k, s);
    sync(&h);
  })
}

int dfmi_set_mass(dfmi_mg* mg, const double* me) {
  DFMI_GUARD_MG({
    if (!me) throw std::invalid_argument("null mass matrix");
    Handle& h = mg->h;
    double mb[24][24];
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        const int qi = kRefToBits[i / 3] * 3 + i % 3, qj = kRefToBits[j / 3] * 3 + j % 3;
        mb[qi][qj] = me[i * 24 + j];
      }
    // Me = M8 (x) I3 with M8 diagonal in the corner-Hadamard basis: D = H M8 H^T, M8 = H^T D H / 64
    double big = 0.0;
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) big = std::max(big, std::fabs(mb[i][j]));
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        const double expect = (i % 3 == j % 3) ? mb[3 * (i / 3)][3 * (j / 3)] : 0.0;
        if (std::fabs(mb[i][j] - expect) > 1e-12 * big)
          throw std::invalid_argument("element mass is not the same scalar matrix for each component");
      }
    double d[8][8];
    for (int m = 0; m < 8; ++m)
      for (int n = 0; n < 8; ++n) {
        double acc = 0.0;
        for (int q = 0; q < 8; ++q)
          for (int r = 0; r < 8; ++r) {
            const double hm = (__builtin_popcount(m & q) & 1) ? -1.0 : 1.0, hn = (__builtin_popcount(n & r) & 1) ? -1.0 : 1.0;
            acc += hm * mb[3 * q][3 * r] * hn;
          }
        d[m][n] = acc;
      }
    for (int m = 0; m < 8; ++m)
      for (int n = 0; n < 8; ++n)
        if (m != n && std::fabs(d[m][n]) > 1e-10 * 64 * big)
          throw std::invalid_argument("element mass lacks the mirror symmetry of an axis-aligned brick");
    for (int m = 0; m < 8; ++m) h.mdiag[m] = d[m][m] / 64.0;
    h.have_mass = true;
  })
}

int dfmi_mass_apply(dfmi_mg* mg, int32_t slot, const double* mass, const double* lumped, const double* x, double* y,
                    int32_t k) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (!h.have_mass) throw std::runtime_error("dfmi_set_mass must precede mass_apply");
    if (k < 1 || !mass || !x || !y) throw std::invalid_argument("mass_apply needs k >= 1 and buffers");
    if (slot < 0 || slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
    ensure_workspace(&h, k);
    hipStream_t s = h.stream;
    h.mass_pad.ensure(sizeof(double) * h.pad.elems());
    DevBuf f32;
    f32.ensure(sizeof(float) * h.pad.elems());
    pad_moduli(h.pad, h.ux, h.uy, h.uz, mass, h.mass_pad.as<double>(), f32.as<float>(), s);
    const double* lump = nullptr;
    if (lumped) {
      flat_to_grid(h.pad, h.ux, h.uy, h.uz, lumped, h.tmp.as<double>(), 1, s);
      lump = h.tmp.as<double>();
    }
    flat_to_grid(h.pad, h.ux, h.uy, h.uz, x, h.b.as<double>(), k, s);
    ApplyArgs a{};
    a.k = k;
    a.x = h.b.as<double>();
    a.y = h.r.as<double>();
    a.dinv = lump;
    for (int m = 0; m < 8; ++m) a.mdiag[m] = h.mdiag[m];
    fine_apply<double>(h.pad, h.mass_pad.as<double>(), h.slots[slot].levels[0]->mask.as<uint8_t>(), h.kb, MODE_MASS, a,
                       s, nullptr);
    grid_to_flat(h.pad, h.ux, h.uy, h.uz, h.r.as<double>(), y, k, s);
    sync(&h);
  })
}

int dfmi_mask(dfmi_mg* mg, int32_t slot, double* x, int32_t k) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (k < 1 || !x) throw std::invalid_argument("mask needs k >= 1 and a buffer");
    if (slot < 0 || slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
    ensure_workspace(&h, k);
    hipStream_t s = h.stream;
    flat_to_grid(h.pad, h.ux, h.uy, h.uz, x, h.b.as<double>(), k, s);
    mask_vector(h.slots[slot].levels[0]->mask.as<uint8_t>(), h.b.as<double>(), h.pad.nodes(), k, s);
    grid_to_flat(h.pad, h.ux, h.uy, h.uz, h.b.as<double>(), x, k, s);
    sync(&h);
  })
}

int dfmi_element_energy(dfmi_mg* mg, const double* u, int32_t k, double* energy) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    ensure_workspace(&h, k);
    flat_to_grid(h.pad, h.ux, h.uy, h.uz, u, h.b.as<double>(), k, h.stream);
    fine_element_energy(h.pad, h.ux, h.uy, h.uz, h.b.as<double>(), h.kb, energy, k, h.stream);
    sync(&h);
  })
}

int dfmi_compliance_sensitivity(dfmi_mg* mg, const double* u, const double* f, int32_t k, const double* dmoduli,
                                double factor, double* compliance, double* sensitivity) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    ensure_workspace(&h, k);
    const long long ne = (long long)h.ux * h.uy * h.uz;
    const long long ndof = 3LL * (h.ux + 1) * (h.uy + 1) * (h.uz + 1);
    DevBuf energy;
    energy.ensure(sizeof(double) * (size_t)k * ne);
    flat_to_grid(h.pad, h.ux, h.uy, h.uz, u, h.b.as<double>(), k, h.stream);
    fine_element_energy(h.pad, h.ux, h.uy, h.uz, h.b.as<double>(), h.kb, energy.as<double>(), k, h.stream);
    compliance_sensitivity(u, f, ndof, k, energy.as<double>(), ne, dmoduli, factor, compliance, sensitivity,
                           h.partial.as<double>(), h.stream);
    sync(&h);
  })
}

int dfmi_modal_sensitivity(dfmi_mg* mg, const double* u, int32_t k, const double* weight, const double* value,
                           const double* stiffness_slope, const double* mass_slope, double target, double* out) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (!h.have_mass) throw std::runtime_error("dfmi_set_mass must precede modal_sensitivity");
    if (!u || !weight || !value || !stiffness_slope || !mass_slope || !out) throw std::invalid_argument("null buffer");
    if (!(target > 0)) throw std::invalid_argument("target must be positive");
    ensure_workspace(&h, k);
    flat_to_grid(h.pad, h.ux, h.uy, h.uz, u, h.b.as<double>(), k, h.stream);
    fine_modal_sensitivity(h.pad, h.ux, h.uy, h.uz, h.b.as<double>(), k, h.kb, h.mdiag, weight, value,
                           stiffness_slope, mass_slope, target, out, h.stream);
    sync(&h);
  })
}

int dfmi_level_apply(dfmi_mg* mg, int32_t slot, int32_t level, const double* x, double* y, int32_t k) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (slot < 0 || slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
    Hierarchy& H = h.slots[slot];
    if (level < 0 || level >= (int)H.levels.size()) throw std::invalid_argument("level out of range");
    Level& L = *H.levels[level];
    ApplyArgs a{};
    a.x = x;
    a.y = y;
    a.k = k;
    if (level == 0)
      fine_apply<double>(L.g, h.m64.as<double>(), L.mask.as<uint8_t>(), h.kb, MODE_APPLY, a, h.stream, nullptr);
    else
      stencil_apply<double>(L.g, L.st64.as<double>(), L.mask.as<uint8_t>(), MODE_APPLY, a, column_major(L.g),
                            h.stream, nullptr);
    sync(&h);
  })
}

int dfmi_precondition(dfmi_mg* mg, int32_t slot, const double* r, double* z, int32_t k) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (slot < 0 || slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
    Hierarchy& H = h.slots[slot];
    ensure_workspace(&h, k);
    std::vector<int> ones(k, 1);
    DFMI_HIP(hipMemcpyAsync(h.active.p, ones.data(), sizeof(int) * k, hipMemcpyHostToDevice, h.stream));
    flat_to_grid(h.pad, h.ux, h.uy, h.uz, r, h.tmp.as<double>(), k, h.stream);
    precond_any(&h, H, h.tmp.as<double>(), h.zn.as<double>(), k, h.active.as<int>());
    grid_to_flat(h.pad, h.ux, h.uy, h.uz, h.zn.as<double>(), z, k, h.stream);
    sync(&h);
  })
}

int dfmi_transfer(dfmi_mg* mg, int32_t slot, int32_t level, int32_t direction, const double* in, double* out,
                  int32_t k) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (slot < 0 || slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
    Hierarchy& H = h.slots[slot];
    if (level < 0 || level + 1 >= (int)H.levels.size()) throw std::invalid_argument("level out of range");
    Level& F = *H.levels[level];
    Level& C = *H.levels[level + 1];
    if (direction == 0) {
      restrict_full<double>(F.g, in, column_major(F.g), C.g, C.mask.as<uint8_t>(), out, column_major(C.g), k,
                            nullptr, h.stream);
    } else {
      fill_zero<double>(out, (long long)k * col(F.g), h.stream);
      prolong_add<double>(C.g, in, column_major(C.g), F.g, F.mask.as<uint8_t>(), out, column_major(F.g), k,
                          nullptr, h.stream);
    }
    sync(&h);
  })
}

int dfmi_time_apply(dfmi_mg* mg, int32_t slot, int32_t level, int32_t precision, int32_t k, int32_t reps,
                    double* ms_per_launch) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (slot < 0 || slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
    if (k < 1 || reps < 1) throw std::invalid_argument("k and reps must be positive");
    Hierarchy& H = h.slots[slot];
    if (level < 0 || level >= (int)H.levels.size() - 1) throw std::invalid_argument("level out of range");
    Level& L = *H.levels[level];
    const long long n = (long long)k * col(L.g);
    const size_t el = precision == 0 ? sizeof(double) : sizeof(float);
    DevBuf xb, yb;
    xb.ensure(el * n);
    yb.ensure(el * n);
    if (precision == 0) fill_zero<double>(xb.as<double>(), n, h.stream);
    else fill_zero<float>(xb.as<float>(), n, h.stream);
    ApplyArgs a{};
    a.x = xb.p;
    a.y = yb.p;
    a.k = k;
    a.premasked = 1;  // the solver's form of the fine kernel (the zero operand is masked)
    auto launch = [&]() {
      if (level > 0 && precision == 0)
        stencil_apply<double>(L.g, L.st64.as<double>(), L.mask.as<uint8_t>(), MODE_APPLY, a, interleaved(k), h.stream,
                              nullptr);
      else if (level > 0)
        stencil_apply<float>(L.g, L.st32.as<float>(), L.mask.as<uint8_t>(), MODE_APPLY, a, interleaved(k), h.stream,
                             nullptr);
      else if (precision == 0)
        fine_apply<double>(h.pad, h.m64.as<double>(), H.levels[0]->mask.as<uint8_t>(), h.kb, MODE_APPLY, a, h.stream,
                           nullptr);
      else if (precision == 2)
        fine_apply_mfma(h.pad, h.m32.as<float>(), H.levels[0]->mask.as<uint8_t>(), h.atab.as<unsigned>(), MODE_APPLY, a,
                        fine_tile(&h, k), h.stream);
      else
        fine_apply<float>(h.pad, h.m32.as<float>(), H.levels[0]->mask.as<uint8_t>(), h.kb, MODE_APPLY, a, h.stream,
                          nullptr);
    };
    launch();
    hipEvent_t e0, e1;
    DFMI_HIP(hipEventCreate(&e0));
    DFMI_HIP(hipEventCreate(&e1));
    DFMI_HIP(hipEventRecord(e0, h.stream));
    for (int i = 0; i < reps; ++i) launch();
    DFMI_HIP(hipEventRecord(e1, h.stream));
    DFMI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DFMI_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_launch = ms / reps;
  })
}

int dfmi_tune_fine(dfmi_mg* mg, int32_t slot, int32_t k, int32_t tile[2], double* ms_per_launch) {
  DFMI_GUARD_MG({
    Handle& h = mg->h;
    if (slot < 0 || slot >= 8 || !h.slots[slot].built) throw std::invalid_argument("slot not built");
    if (k < 1) throw std::invalid_argument("k must be positive");
    if (!fine_mfma_supported(h.pad, MODE_CHEB)) throw std::invalid_argument("grid not supported on the matrix cores");
    Hierarchy& H = h.slots[slot];
    if (tile && tile[0] > 0) {  // set this tile for k without timing
      if ((tile[0] != 8 && tile[0] != 16) || tile[1] < 1) throw std::invalid_argument("tile must be {8 or 16, >= 1}");
      h.mfma_tiles[k] = FineMfmaTile{tile[0], tile[1]};
      if (ms_per_launch) *ms_per_launch = 0.0;
      return 0;
    }
    ensure_level_vectors(&h, H, k);
    double ms = 0.0;
    const FineMfmaTile best = tune_fine(&h, H, k, &ms);
    if (tile) tile[0] = best.tye, tile[1] = best.xchunk;
    if (ms_per_launch) *ms_per_launch = ms;
  })
}

int dfmi_diagnostics(dfmi_mg* mg, int32_t slot, double* lambda_max, int32_t* coarsest_dofs) {
  DFMI_GUARD_MG({
    if (slot < 0 || slot >= 8) throw std::invalid_argument("slot must be in [0, 8)");
    Hierarchy& H = mg->h.slots[slot];
    if (!H.built) throw std::invalid_argument("slot not built");
    for (size_t l = 0; l + 1 < H.levels.size(); ++l) lambda_max[l] = H.levels[l]->lambda_max;
    *coarsest_dofs = H.coarsest_n;
  })
}

int dfmi_history(dfmi_mg* mg, double* out, int32_t capacity, int32_t* count) {
  DFMI_GUARD_MG({
    const auto& hist = mg->h.history;
    if (count) *count = (int32_t)hist.size();
    if (out)
      for (int i = 0; i < std::min<int>(capacity, (int)hist.size()); ++i) out[i] = hist[i];
  })
}

int dfmi_kernel_check(dfmi_mg* mg, int32_t slot, double* out, int32_t* count) {
  DFMI_GUARD_MG({
    if (slot < 0 || slot >= 8 || !mg->h.slots[slot].built) throw std::invalid_argument("slot not built");
    const auto& kc = mg->h.slots[slot].kernel_check;
    if (count) *count = (int32_t)kc.size();
    if (out)
      for (size_t i = 0; i < kc.size(); ++i) out[i] = kc[i];
  })
}


// ---- learned coarse correction ----------------------------------------------------------------
int dfmi_net_create(dfmi_net** out) {
  DFMI_GUARD({
    if (!out) throw std::invalid_argument("null output handle");
    auto* n = new dfmi_net();
    n->impl = net_new();
    *out = n;
  })
}

int dfmi_net_destroy(dfmi_net* net) {
  DFMI_GUARD({
    if (net) {
      net_delete(net->impl);
      delete net;
    }
  })
}

int dfmi_net_set_stream(dfmi_net* net, void* stream) {
  DFMI_GUARD({ net_set_stream(net->impl, (hipStream_t)stream); })
}

int dfmi_net_param_count(dfmi_net* net, int64_t* count) { DFMI_GUARD({ *count = net_param_count(net->impl); }) }

int dfmi_net_set_params(dfmi_net* net, const float* params) {
  DFMI_GUARD({
    if (!params) throw std::invalid_argument("null parameter buffer");
    net_set_params(net->impl, params);
  })
}

int dfmi_net_get(dfmi_net* net, int32_t which, float* out) {
  DFMI_GUARD({
    if (!out || (which != 0 && which != 1)) throw std::invalid_argument("which must be 0 (weights) or 1 (gradients)");
    net_get(net->impl, which, out);
  })
}

int dfmi_net_forward(dfmi_net* net, const float* x, int32_t n, const int32_t dims[3], float* u) {
  DFMI_GUARD({
    if (!x || !u || n < 1 || dims[0] < 1 || dims[1] < 1 || dims[2] < 1)
      throw std::invalid_argument("forward needs n >= 1, positive dims and buffers");
    const NetDims d{dims[0], dims[1], dims[2]};
    net_forward(net->impl, n, d, x);
    net_copy_output(net->impl, u, (long long)n * 3 * d.n());
  })
}

int dfmi_net_backward(dfmi_net* net, const float* du) {
  DFMI_GUARD({
    if (!du) throw std::invalid_argument("null gradient");
    net_backward(net->impl, du);
  })
}

int dfmi_net_adam(dfmi_net* net, double lr, double beta1, double beta2, double eps) {
  DFMI_GUARD({ net_adam(net->impl, lr, beta1, beta2, eps); })
}

int dfmi_net_train_step(dfmi_net* net, dfmi_mg* mg, int32_t slot, const float* g, int32_t n, double lr,
                        double* loss) {
  DFMI_GUARD_MG({
    if (slot < 0 || slot >= 8 || !mg->h.slots[slot].built) throw std::invalid_argument("hierarchy slot not built");
    Hierarchy& H = mg->h.slots[slot];
    if (H.levels.size() < 2) throw std::invalid_argument("training needs at least two levels");
    if (!g || n < 1) throw std::invalid_argument("train step needs n >= 1 samples");
    Level& C = *H.levels[1];
    hipStream_t s = mg->h.stream;
    net_set_stream(net->impl, s);
    auto apply = [&](const float* x, float* y) {
      ApplyArgs a{};
      a.x = x;
      a.y = y;
      a.k = n;
      stencil_apply<float>(C.g, C.st32.as<float>(), C.mask.as<uint8_t>(), MODE_APPLY, a, column_major(C.g), s,
                           nullptr);
    };
    const double l = net_train_step(net->impl, C.g, g, n, C.dinv32.as<float>(), lr, apply);
    if (loss) *loss = l;
  })
}

int dfmi_attach_net(dfmi_mg* mg, dfmi_net* net) {
  DFMI_GUARD_MG({
    if (net && mg->h.st.precision != 1) throw std::invalid_argument("the learned correction needs precision fp32");
    mg->h.net = net ? net->impl : nullptr;
    if (net) net_invalidate_feature(net->impl);
  })
}

int dfmi_harvest(dfmi_mg* mg, float* buffer, int64_t capacity) {
  DFMI_GUARD_MG({
    if (capacity < 0 || (capacity > 0 && !buffer)) throw std::invalid_argument("harvest needs a buffer");
    mg->h.harvest = capacity > 0 ? buffer : nullptr;
    mg->h.harvest_cap = capacity;
    mg->h.harvest_count = 0;
  })
}

int dfmi_harvest_count(dfmi_mg* mg, int64_t* count) { DFMI_GUARD_MG({ *count = mg->h.harvest_count; }) }
}  // extern "C"
