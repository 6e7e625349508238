// libdgfem.so: DG-FEM right-hand sides (wave, Euler, Maxwell; affine
// tetrahedra and triangles, P1-P4) for gfx950, behind the C ABI declared in
// include/dgfem.h.
//
// One fused kernel per (equation, order) evaluates the whole RHS of a batch of
// elements: the volume contraction with the reference derivative matrices, the
// trace gather through the exterior-node map, the numerical flux and the LIFT
// (face mass + inverse mass) application.  State is read once from HBM (plus
// the neighbours' face nodes, which mostly hit L2 / MALL) and the RHS written
// once.  Operator forms follow the reference optimizer's canonical einsums:
//   weak divergence  xre,rij,xej->ei   (mass_inverse_fuser.py:39-56)
//   weak gradient    re,rij,ej->ei     (mass_inverse_fuser.py:57-74)
//   face mass/LIFT   ifj,fe,fej->ei    (mass_inverse_fuser.py:21-38)
// Two tetrahedral variants: the MFMA kernel (rhs_mfma.h) and an LDS-staged
// one-thread-per-node kernel (rhs_lds.h); the measured-faster one is the
// default per equation.  Triangles run the LDS-staged kernel of rhs_2d.h.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <array>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "dgfem.h"
#include "dg_common.h"
#include "rhs_2d.h"
#include "rhs_lds.h"
#include "rhs_mfma.h"
#include "rhs_mixed.h"
#include "rhs_adjoint.h"
#include "rhs_adjoint_mfma.h"

// Workspace of dg_rhs_vjp_noalloc, for callers that cannot allocate: one
// copy per GPU the library runs on, held until the process ends.  Host code
// never names it: its device address comes from hipGetSymbolAddress, looked
// up once per device when a plan with the adjoint is created there.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

using namespace dg;

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define DG_HIP_CHECK(expr)                                                   \
  do {                                                                       \
    hipError_t err_ = (expr);                                                \
    if (err_ != hipSuccess)                                                  \
      return fail(DG_ERR_HIP, std::string(#expr) + ": " +                    \
                                  hipGetErrorString(err_));                  \
  } while (0)

// This device's address of `scratch`, looked up on the first call per
// device (with that device current) and kept.
int scratch_address(int device, void** out) {
  static std::mutex mu;
  static std::vector<void*> addr;
  std::lock_guard<std::mutex> lock(mu);
  if (device < 0) return fail(DG_ERR_HIP, "scratch: invalid device");
  if ((size_t)device >= addr.size()) addr.resize((size_t)device + 1, nullptr);
  if (!addr[(size_t)device]) {
    void* ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !ptr)
      return fail(DG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                  hipGetErrorString(err));
    addr[(size_t)device] = ptr;
  }
  *out = addr[(size_t)device];
  return DG_OK;
}

// {{{ LSERK stage

struct LserkArgs {
  OutPtrs q;
  OutPtrs res;
  InPtrs rhs;
  int64_t n;
  int nf;
  double a, b, dt;
};

__global__ __launch_bounds__(kBlock) void lserk_stage_kernel(LserkArgs s) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int f = 0; f < s.nf; ++f) {
    double* q = s.q.p[f];
    double* res = s.res.p[f];
    const double* rhs = s.rhs.p[f];
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < s.n; k += stride) {
      // a == 0 (the first stage) does not read res, so it may start uninitialised
      const double r = (s.a != 0.0 ? s.a * res[k] : 0.0) + s.dt * rhs[k];
      res[k] = r;
      q[k] += s.b * r;
    }
  }
}

// Reverse sweep of one stage: q_bar += dt*w (w optional); s = a*s + b*q_bar
// (s not read when a == 0).  float64, or float32 for the FP32 adjoint.
template <typename T>
struct LserkAdjArgs {
  T* qbar[kMaxFields];
  T* s[kMaxFields];
  const T* w[kMaxFields];
  int64_t n;
  int nf;
  double a, b, dt;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void lserk_adjoint_kernel(LserkAdjArgs<T> s) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const T a = (T)s.a, b = (T)s.b, dt = (T)s.dt;
  for (int f = 0; f < s.nf; ++f) {
    T* qb = s.qbar[f];
    T* sv = s.s[f];
    const T* w = s.w[f];
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < s.n; k += stride) {
      T q = qb[k];
      if (w) {
        q += dt * w[k];
        qb[k] = q;
      }
      sv[k] = (s.a != 0.0 ? a * sv[k] : (T)0) + b * q;
    }
  }
}

// }}}

using RhsKernel = void (*)(RhsArgs);

// A kernel and its launch shape: threads per block and elements per batch.
struct KernelInfo {
  RhsKernel fn = nullptr;
  int threads = 0;
  int batch_elems = 0;
};

template <int EQ, int ORDER>
KernelInfo lds_info() {
  return {rhs_lds_kernel<EQ, ORDER>, kBlock, kBlock / np_of(ORDER)};
}

template <int EQ, int ORDER>
KernelInfo mfma_info() {
  return {rhs_mfma_kernel<EQ, ORDER>, 64 * mfma_waves_of(EQ, ORDER),
          16 * mfma_waves_of(EQ, ORDER)};
}

#define DG_DEFINE_PICK(NAME, INFO)                                           \
  template <int EQ>                                                          \
  KernelInfo NAME##_order(int order) {                                       \
    switch (order) {                                                         \
      case 1: return INFO<EQ, 1>();                                          \
      case 2: return INFO<EQ, 2>();                                          \
      case 3: return INFO<EQ, 3>();                                          \
      case 4: return INFO<EQ, 4>();                                          \
      default: return KernelInfo{};                                          \
    }                                                                        \
  }                                                                          \
  KernelInfo NAME(int eq, int order) {                                       \
    switch (eq) {                                                            \
      case DG_EQ_WAVE: return NAME##_order<DG_EQ_WAVE>(order);               \
      case DG_EQ_EULER: return NAME##_order<DG_EQ_EULER>(order);             \
      case DG_EQ_MAXWELL: return NAME##_order<DG_EQ_MAXWELL>(order);         \
      default: return KernelInfo{};                                          \
    }                                                                        \
  }

DG_DEFINE_PICK(pick_lds, lds_info)
DG_DEFINE_PICK(pick_mfma, mfma_info)

// The MFMA kernel with the fused mass post-multiplication, for the adjoint of
// the linear operators (wave, Maxwell).
template <int EQ>
RhsKernel pick_mfma_post_order(int order) {
  switch (order) {
    case 1: return rhs_mfma_kernel<EQ, 1, true>;
    case 2: return rhs_mfma_kernel<EQ, 2, true>;
    case 3: return rhs_mfma_kernel<EQ, 3, true>;
    case 4: return rhs_mfma_kernel<EQ, 4, true>;
    default: return nullptr;
  }
}
// The MFMA kernel with the LSERK4 stage fused into its output.
template <int EQ>
RhsKernel pick_mfma_stage_order(int order) {
  switch (order) {
    case 1: return rhs_mfma_kernel<EQ, 1, false, true>;
    case 2: return rhs_mfma_kernel<EQ, 2, false, true>;
    case 3: return rhs_mfma_kernel<EQ, 3, false, true>;
    case 4: return rhs_mfma_kernel<EQ, 4, false, true>;
    default: return nullptr;
  }
}
RhsKernel pick_mfma_stage(int eq, int order) {
  switch (eq) {
    case DG_EQ_WAVE: return pick_mfma_stage_order<DG_EQ_WAVE>(order);
    case DG_EQ_EULER: return pick_mfma_stage_order<DG_EQ_EULER>(order);
    case DG_EQ_MAXWELL: return pick_mfma_stage_order<DG_EQ_MAXWELL>(order);
    default: return nullptr;
  }
}

RhsKernel pick_mfma_post(int eq, int order) {
  switch (eq) {
    case DG_EQ_WAVE: return pick_mfma_post_order<DG_EQ_WAVE>(order);
    case DG_EQ_MAXWELL: return pick_mfma_post_order<DG_EQ_MAXWELL>(order);
    default: return nullptr;
  }
}

template <typename T>
using TriKernel = void (*)(Rhs2DArgs<T>);

template <int EQ, typename T>
TriKernel<T> pick_2d_order(int order) {
  switch (order) {
    case 1: return rhs_2d_kernel<EQ, 1, T>;
    case 2: return rhs_2d_kernel<EQ, 2, T>;
    case 3: return rhs_2d_kernel<EQ, 3, T>;
    case 4: return rhs_2d_kernel<EQ, 4, T>;
    default: return nullptr;
  }
}

template <typename T>
TriKernel<T> pick_2d(int eq, int order) {
  switch (eq) {
    case DG_EQ_WAVE: return pick_2d_order<DG_EQ_WAVE, T>(order);
    case DG_EQ_EULER: return pick_2d_order<DG_EQ_EULER, T>(order);
    case DG_EQ_MAXWELL: return pick_2d_order<DG_EQ_MAXWELL, T>(order);
    default: return nullptr;
  }
}

using MixedKernel = void (*)(MixedArgs);

template <int EQ, int PREC, bool STAGE>
MixedKernel pick_mixed_order(int order) {
  switch (order) {
    case 1: return rhs_mixed_kernel<EQ, PREC, 1, STAGE>;
    case 2: return rhs_mixed_kernel<EQ, PREC, 2, STAGE>;
    case 3: return rhs_mixed_kernel<EQ, PREC, 3, STAGE>;
    case 4: return rhs_mixed_kernel<EQ, PREC, 4, STAGE>;
    default: return nullptr;
  }
}

template <int PREC, bool STAGE>
MixedKernel pick_mixed_prec(int eq, int order) {
  switch (eq) {
    case DG_EQ_WAVE: return pick_mixed_order<DG_EQ_WAVE, PREC, STAGE>(order);
    case DG_EQ_EULER: return pick_mixed_order<DG_EQ_EULER, PREC, STAGE>(order);
    case DG_EQ_MAXWELL: return pick_mixed_order<DG_EQ_MAXWELL, PREC, STAGE>(order);
    default: return nullptr;
  }
}

// The mixed RHS kernel, or (stage) its variant with the LSERK4 stage fused.
MixedKernel pick_mixed(int prec, int eq, int order, bool stage = false) {
  if (prec == DG_PREC_FP32)
    return stage ? pick_mixed_prec<DG_PREC_FP32, true>(eq, order)
                 : pick_mixed_prec<DG_PREC_FP32, false>(eq, order);
  if (prec == DG_PREC_BF16X3)
    return stage ? pick_mixed_prec<DG_PREC_BF16X3, true>(eq, order)
                 : pick_mixed_prec<DG_PREC_BF16X3, false>(eq, order);
  return nullptr;
}

// The FP64 kernel "auto" picks, from measurements on MI355X at 3-4M DOFs
// per field: the MFMA kernel.  It leads the LDS-staged kernel by 10-60%
// (wave P3: 112 vs 173 us); at Maxwell P2 it does so only with the direct
// flux, which frees the LDS for three waves per SIMD (79 against 88 us; with
// the chunked producer it held one wave per SIMD and was 11% slower).
// Whether every element's face geometry is the affine image of the reference
// element's: Fscale_f n_f = rx^T c_f with c_f = ref_face_coef (relative
// 1e-10), which the MFMA kernel's reference-normal LIFT relies on.
bool geometry_fits_ref_normals(int64_t k, const double* geo) {
  for (int64_t e = 0; e < k; ++e) {
    const double* g = geo + e * kGeo;
    for (int f = 0; f < kFaces; ++f) {
      const double fs = g[9 + 4 * f + 3];
      for (int x = 0; x < 3; ++x) {
        double v = 0.0, mag = 0.0;
        for (int r = 0; r < 3; ++r) {
          v += g[3 * x + r] * ref_face_coef(f, r);
          mag += std::fabs(g[3 * x + r]);
        }
        const double want = fs * g[9 + 4 * f + x];
        if (!(std::fabs(v - want) <= 1e-10 * (mag + std::fabs(want)))) return false;
      }
    }
  }
  return true;
}

int auto_variant(int equation, int order) {
  (void)equation;
  (void)order;
  return DG_KERNEL_MFMA;
}

// {{{ locality schedule

// Work order for units of `unit` consecutive elements (see Schedule in
// dg_common.h).  With element centroids, the units are split into 8 parts of
// equal count by recursive coordinate bisection (longest axis at the median).
// Each part is swept in kSlabsPerPart slabs across its longest axis; a slab
// is walked in rows of kSlabRow units along the second-longest axis, and a
// row along the third.  Every XCD thus works on a compact eighth of the mesh,
// and the data a slab gathers from its predecessor was read a slab earlier:
// at most a quarter of the part back, which the shared last-level cache
// still holds when the state exceeds it.  Without centroids the units keep
// their storage order and the parts are 8 contiguous runs.
constexpr int kSlabsPerPart = 4;
constexpr int kSlabRow = 16;

void build_schedule(int64_t nelem, int unit, const double* centroids,
                    std::vector<int32_t>& order, std::array<int32_t, 9>& part) {
  const int64_t nunits = (nelem + unit - 1) / unit;
  order.resize(nunits);
  std::iota(order.begin(), order.end(), 0);
  for (int x = 0; x <= 8; ++x) part[x] = (int32_t)(nunits * x / 8);
  if (!centroids) return;
  std::vector<std::array<double, 3>> c(nunits, {0.0, 0.0, 0.0});
  for (int64_t u = 0; u < nunits; ++u) {
    const int64_t e0 = u * unit, e1 = std::min<int64_t>(nelem, e0 + unit);
    for (int64_t e = e0; e < e1; ++e)
      for (int d = 0; d < 3; ++d) c[u][d] += centroids[e * 3 + d];
    for (int d = 0; d < 3; ++d) c[u][d] /= (double)(e1 - e0);
  }
  auto bbox = [&](int64_t b, int64_t e, double* lo, double* hi) {
    for (int d = 0; d < 3; ++d) {
      lo[d] = 1e300;
      hi[d] = -1e300;
    }
    for (int64_t k = b; k < e; ++k)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], c[order[k]][d]);
        hi[d] = std::max(hi[d], c[order[k]][d]);
      }
  };
  // bisection into parts [part[x], part[x+1]) of 8 equal counts
  struct Range { int64_t b, e; int p0, np; };
  std::vector<Range> stack{{0, nunits, 0, 8}};
  while (!stack.empty()) {
    const Range r = stack.back();
    stack.pop_back();
    double lo[3], hi[3];
    bbox(r.b, r.e, lo, hi);
    if (r.np == 1) {
      int ax[3] = {0, 1, 2};
      std::stable_sort(ax, ax + 3, [&](int x, int y) { return hi[x] - lo[x] > hi[y] - lo[y]; });
      auto along = [&](int axis) {
        return [&c, axis](int32_t x, int32_t y) {
          return c[x][axis] < c[y][axis] || (c[x][axis] == c[y][axis] && x < y);
        };
      };
      const int64_t slab = std::max<int64_t>(1, (r.e - r.b + kSlabsPerPart - 1) / kSlabsPerPart);
      std::sort(order.begin() + r.b, order.begin() + r.e, along(ax[0]));
      for (int64_t sb = r.b; sb < r.e; sb += slab) {
        const int64_t se = std::min<int64_t>(r.e, sb + slab);
        std::sort(order.begin() + sb, order.begin() + se, along(ax[1]));
        for (int64_t rb = sb; rb < se; rb += kSlabRow)
          std::sort(order.begin() + rb, order.begin() + std::min<int64_t>(se, rb + kSlabRow),
                    along(ax[2]));
      }
      continue;
    }
    int axis = 0;
    for (int d = 1; d < 3; ++d)
      if (hi[d] - lo[d] > hi[axis] - lo[axis]) axis = d;
    const int half = r.np / 2;
    const int64_t mid = part[r.p0 + half];
    if (mid > r.b && mid < r.e)
      std::nth_element(order.begin() + r.b, order.begin() + mid, order.begin() + r.e,
                       [&](int32_t x, int32_t y) {
                         return c[x][axis] < c[y][axis] || (c[x][axis] == c[y][axis] && x < y);
                       });
    stack.push_back({r.b, mid, r.p0, half});
    stack.push_back({mid, r.e, r.p0 + half, r.np - half});
  }
}

// }}}

// {{{ adjoint kernels and their host-side data

template <typename T>
using VjpKernel = void (*)(VjpArgs<T>);
template <typename T>
using MassKernel = void (*)(MassArgs<T>);

template <int EQ, typename T, int DIM>
VjpKernel<T> pick_vjp_order(int order) {
  switch (order) {
    case 1: return rhs_vjp_kernel<EQ, 1, T, DIM>;
    case 2: return rhs_vjp_kernel<EQ, 2, T, DIM>;
    case 3: return rhs_vjp_kernel<EQ, 3, T, DIM>;
    case 4: return rhs_vjp_kernel<EQ, 4, T, DIM>;
    default: return nullptr;
  }
}

// The FP64 adjoint of the linear operators on tetrahedra runs on the forward
// kernel (rhs_vjp_impl), so there they have FP32 adjoint kernels only.
template <typename T, int DIM>
VjpKernel<T> pick_vjp_dim(int eq, int order) {
  constexpr bool kernel_for_linear = sizeof(T) == 4 || DIM == 2;
  switch (eq) {
    case DG_EQ_WAVE:
      if constexpr (kernel_for_linear) return pick_vjp_order<DG_EQ_WAVE, T, DIM>(order);
      return nullptr;
    case DG_EQ_EULER: return pick_vjp_order<DG_EQ_EULER, T, DIM>(order);
    case DG_EQ_MAXWELL:
      if constexpr (kernel_for_linear) return pick_vjp_order<DG_EQ_MAXWELL, T, DIM>(order);
      return nullptr;
    default: return nullptr;
  }
}

template <typename T>
VjpKernel<T> pick_vjp(int eq, int order, int dim) {
  return dim == 2 ? pick_vjp_dim<T, 2>(eq, order) : pick_vjp_dim<T, 3>(eq, order);
}

// The Euler adjoint on the matrix cores (FP64, tetrahedra).
using VjpMfmaKernel = void (*)(VjpMfmaArgs);
VjpMfmaKernel pick_vjp_mfma(int order) {
  switch (order) {
    case 1: return rhs_vjp_mfma_kernel<1>;
    case 2: return rhs_vjp_mfma_kernel<2>;
    case 3: return rhs_vjp_mfma_kernel<3>;
    case 4: return rhs_vjp_mfma_kernel<4>;
    default: return nullptr;
  }
}

// A tiles of rhs_vjp_mfma_kernel (operand layouts in rhs_mfma.h; m(l) = l % 4
// on an output block of at most 4 rows, else l % 16):
//   volume: vol[(r * MB + b) * KS + ks][l] = A_r[16 b + m(l)][(l / 16) * KS + ks],
//           A_r[i][j] = amat_t[r][j][i];
//   faces:  face[f * KSF + ks][l] = Mf_f[b = l / 16 + 4 ks][a = m(l)] (the
//           contraction psi[a] = sum_b Mf_f[b][a] mu[b]);
//   own:    own[f * KS + ks][l] = P_f[a = m(l)][j = (l / 16) * KS + ks], P_f[a][j] =
//           Mf_f[b][a] where fmask_f[b] = j: the same contraction on volume nodes.
// The exterior map per tile, lane and face node: ext_v[t][l][f * KSF + v] =
// ext_map[e][f * Nfp + a] for a = l / 16 + 4 v < Nfp (element e = 16 t + l % 16,
// the last one repeated past the end), else -1; with `intile`
// (vjp_mfma_intile), a neighbour inside the tile as kInTile | local element
// << 8 | node.
void build_vjp_mfma_tables(int64_t k, int np, int nfp, const double* amat_t, const double* fm,
                           const int32_t* fmask, const int32_t* emap, bool intile,
                           std::vector<double>& vol,
                           std::vector<double>& face, std::vector<double>& own,
                           std::vector<int32_t>& ext_v) {
  const int mb = (np + 15) / 16, ks_n = (np + 3) / 4, ksf = (nfp + 3) / 4;
  const bool shf = nfp <= 4;
  vol.assign((size_t)3 * mb * ks_n * 64, 0.0);
  for (int r = 0; r < 3; ++r)
    for (int b = 0; b < mb; ++b)
      for (int ks = 0; ks < ks_n; ++ks)
        for (int l = 0; l < 64; ++l) {
          const int m = 16 * b + ((b == mb - 1 && short_last_block(np)) ? (l & 3) : (l & 15));
          const int j = (l >> 4) * ks_n + ks;
          if (m < np && j < np)
            vol[(((size_t)r * mb + b) * ks_n + ks) * 64 + l] = amat_t[((size_t)r * np + j) * np + m];
        }
  face.assign((size_t)kFaces * ksf * 64, 0.0);
  own.assign((size_t)kFaces * ks_n * 64, 0.0);
  for (int f = 0; f < kFaces; ++f) {
    for (int ks = 0; ks < ksf; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int aa = shf ? (l & 3) : (l & 15);
        const int bb = (l >> 4) + 4 * ks;
        if (aa < nfp && bb < nfp)
          face[((size_t)f * ksf + ks) * 64 + l] = fm[((size_t)f * nfp + bb) * nfp + aa];
      }
    for (int ks = 0; ks < ks_n; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int aa = shf ? (l & 3) : (l & 15);
        const int j = (l >> 4) * ks_n + ks;
        if (aa >= nfp || j >= np) continue;
        for (int bb = 0; bb < nfp; ++bb)
          if (fmask[f * nfp + bb] == j)
            own[((size_t)f * ks_n + ks) * 64 + l] += fm[((size_t)f * nfp + bb) * nfp + aa];
      }
  }
  const int64_t ntiles = (k + 15) / 16;
  const int next = kFaces * ksf, nfn = kFaces * nfp;
  ext_v.assign((size_t)ntiles * 64 * next, -1);
  for (int64_t t = 0; t < ntiles; ++t)
    for (int l = 0; l < 64; ++l) {
      const int64_t e = std::min<int64_t>(t * 16 + (l & 15), k - 1);
      for (int f = 0; f < kFaces; ++f)
        for (int v = 0; v < ksf; ++v) {
          const int aa = (l >> 4) + 4 * v;
          if (aa >= nfp) continue;
          int32_t x = emap[e * nfn + f * nfp + aa];
          // neighbours inside the tile: kInTile | local element << 8 | node
          if (intile && x >= 0 && x / np / 16 == t)
            x = kInTile | (int32_t)(((x / np) % 16) << 8) | (x % np);
          ext_v[((size_t)t * 64 + l) * next + f * ksf + v] = x;
        }
    }
}

// Mass-operator kernels by node count (tetrahedra: 4, 10, 20, 35; triangles:
// 3, 6, 10, 15) and field count (3 to 6)
template <typename T, int NP>
MassKernel<T> mass_of_fields(int nf) {
  switch (nf) {
    case 3: return mass_apply_kernel<NP, 3, T>;
    case 4: return mass_apply_kernel<NP, 4, T>;
    case 5: return mass_apply_kernel<NP, 5, T>;
    case 6: return mass_apply_kernel<NP, 6, T>;
    default: return nullptr;
  }
}

template <typename T>
MassKernel<T> pick_mass(int np, int nf) {
  switch (np) {
    case 3: return mass_of_fields<T, 3>(nf);
    case 4: return mass_of_fields<T, 4>(nf);
    case 6: return mass_of_fields<T, 6>(nf);
    case 10: return mass_of_fields<T, 10>(nf);
    case 15: return mass_of_fields<T, 15>(nf);
    case 20: return mass_of_fields<T, 20>(nf);
    case 35: return mass_of_fields<T, 35>(nf);
    default: return nullptr;
  }
}

// Inverse of a small dense matrix (Gauss-Jordan, partial pivoting); false if
// singular.
bool invert(int n, const double* a, std::vector<double>& inv) {
  std::vector<double> m(a, a + (size_t)n * n);
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(m[(size_t)r * n + c]) > std::fabs(m[(size_t)piv * n + c])) piv = r;
    if (m[(size_t)piv * n + c] == 0.0) return false;
    for (int k = 0; k < n; ++k) {
      std::swap(m[(size_t)c * n + k], m[(size_t)piv * n + k]);
      std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
    }
    const double d = 1.0 / m[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      m[(size_t)c * n + k] *= d;
      inv[(size_t)c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = m[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        m[(size_t)r * n + k] -= f * m[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return true;
}

// Host data of the adjoint: per-face blocks of the face mass, the face nodes
// of every volume node, and the neighbour's Fscale across every face.  Fails
// (returns a message) when E couples nodes off its face or the exterior map
// does not pair whole faces.
// nfaces faces per element; geometry rows of geo_stride doubles with each
// face's Fscale at fs_offset + fs_stride * f.
std::string build_adjoint_data(const dg_plan_desc* d, int np, int nfp, int nfaces, int geo_stride,
                               int fs_offset, int fs_stride, std::vector<double>& face_mass,
                               std::vector<int32_t>& node_faces, std::vector<double>& ext_fscale) {
  const int kFaces = nfaces;
  const int nfn = kFaces * nfp;
  const int64_t k = d->nelements;
  const int* fmask = d->face_nodes;
  double emax = 0.0;
  for (int t = 0; t < np * nfn; ++t) emax = std::fmax(emax, std::fabs(d->face_mass[t]));
  std::vector<int> on_face((size_t)kFaces * np, -1);  // face-node index b of node i on face f
  for (int f = 0; f < kFaces; ++f)
    for (int b = 0; b < nfp; ++b) on_face[f * np + fmask[f * nfp + b]] = b;
  face_mass.assign((size_t)kFaces * nfp * nfp, 0.0);
  for (int f = 0; f < kFaces; ++f)
    for (int i = 0; i < np; ++i)
      for (int a = 0; a < nfp; ++a) {
        const double v = d->face_mass[(size_t)i * nfn + f * nfp + a];
        const int b = on_face[f * np + i];
        if (b < 0) {
          if (std::fabs(v) > 1e-12 * emax) return "face_mass couples a node off its face";
        } else {
          face_mass[((size_t)f * nfp + b) * nfp + a] = v;
        }
      }
  node_faces.assign((size_t)np * 3, -1);
  std::vector<int> bits(np, 0);
  for (int f = 0; f < kFaces; ++f)
    for (int b = 0; b < nfp; ++b) {
      const int i = fmask[f * nfp + b];
      int s = 0;
      while (s < 3 && node_faces[i * 3 + s] >= 0) ++s;
      if (s == 3) return "face_nodes: a node lies on more than three faces";
      node_faces[i * 3 + s] = f * nfp + b;
      bits[i] |= 1 << f;
    }
  ext_fscale.assign((size_t)k * kFaces, 0.0);
  for (int64_t e = 0; e < k; ++e)
    for (int f = 0; f < kFaces; ++f) {
      const int32_t* em = d->ext_map + e * nfn + f * nfp;
      if (em[0] < 0) {
        for (int a = 1; a < nfp; ++a)
          if (em[a] >= 0) return "ext_map mixes boundary and interior nodes on one face";
        continue;
      }
      const int64_t e2 = em[0] / np;
      int mask = (1 << kFaces) - 1;
      for (int a = 0; a < nfp; ++a) {
        if (em[a] < 0 || em[a] / np != e2) return "ext_map does not pair whole faces";
        mask &= bits[em[a] % np];
      }
      if (mask == 0 || (mask & (mask - 1)) != 0)
        return "ext_map: exterior nodes of a face do not lie on one face";
      int f2 = 0;
      while (!(mask & (1 << f2))) ++f2;
      ext_fscale[e * kFaces + f] = d->geometry[e2 * geo_stride + fs_offset + fs_stride * f2];
    }
  return "";
}

// }}}

// A tiles of the mixed kernel (layout in rhs_mixed.h): 32-bit words, either
// one float per lane (FP32) or four (hi, lo) bf16 pairs packed in 4 words.
void build_mixed_tiles(int prec, int kgb, int eq, int order, const double* op, const double* lift,
                       std::vector<uint32_t>& op_words, std::vector<uint32_t>& lift_words) {
  const int np = np_of(order), nfp = nfp_of(order);
  const int kg = prec == DG_PREC_BF16X3 ? kgb : 1;
  const int aw = kg;
  const int mb = (np + 15) / 16;
  const int ks_n = ((np + 3) / 4 + kg - 1) / kg * kg;
  const int ksl = (nfp + kg - 1) / kg * kg;
  const int nfn = kFaces * nfp;
  auto pack = [&](std::vector<uint32_t>& dst, size_t base, const double* vals) {
    if (prec == DG_PREC_FP32) {
      const float f = (float)vals[0];
      uint32_t u;
      std::memcpy(&u, &f, 4);
      dst[base] = u;
      return;
    }
    // kg k per lane: words 0 .. kg/2-1 the hi parts, then the lo parts, two per word
    uint16_t hi[8], lo[8];
    for (int i = 0; i < kg; ++i) {
      const float f = (float)vals[i];
      hi[i] = bf16_bits(f);
      lo[i] = bf16_bits(f - bf16_value(hi[i]));
    }
    for (int w = 0; w < kg / 2; ++w) {
      dst[base + w] = hi[2 * w] | ((uint32_t)hi[2 * w + 1] << 16);
      dst[base + kg / 2 + w] = lo[2 * w] | ((uint32_t)lo[2 * w + 1] << 16);
    }
  };
  op_words.assign((size_t)3 * mb * (ks_n / kg) * 64 * aw, 0u);
  lift_words.assign((size_t)mb * (ksl / kg) * 64 * aw, 0u);
  double vals[8];
  for (int r = 0; r < 3; ++r)
    for (int b = 0; b < mb; ++b)
      for (int kq = 0; kq < ks_n / kg; ++kq)
        for (int l = 0; l < 64; ++l) {
          const int m = 16 * b + (l & 15);
          for (int i = 0; i < kg; ++i) {
            const int j = (l >> 4) * ks_n + kq * kg + i;
            vals[i] = (m < np && j < np) ? op[((size_t)r * np + m) * np + j] : 0.0;
          }
          pack(op_words, ((((size_t)r * mb + b) * (ks_n / kg) + kq) * 64 + l) * aw, vals);
        }
  if (mixed_direct_flux_of(eq, prec, order)) {
    // the direct flux's k order: k index i of lane group g in instruction kq
    // holds face-node pair mixed_pair_of(eq, prec, order, kq, g, i)
    const int kqd = mixed_direct_kq_of(eq, prec, order);
    lift_words.assign((size_t)mb * kqd * 64 * aw, 0u);
    for (int b = 0; b < mb; ++b)
      for (int kq = 0; kq < kqd; ++kq)
        for (int l = 0; l < 64; ++l) {
          const int m = 16 * b + (l & 15);
          for (int i = 0; i < kg; ++i) {
            const int fp = mixed_pair_of(eq, prec, order, kq, l >> 4, i);
            vals[i] = (m < np && fp >= 0) ? lift[(size_t)m * nfn + fp] : 0.0;
          }
          pack(lift_words, (((size_t)b * kqd + kq) * 64 + l) * aw, vals);
        }
    return;
  }
  for (int b = 0; b < mb; ++b)
    for (int kq = 0; kq < ksl / kg; ++kq)
      for (int l = 0; l < 64; ++l) {
        const int m = 16 * b + (l & 15);
        for (int i = 0; i < kg; ++i) {
          const int ks = kq * kg + i;
          vals[i] = (m < np && ks < nfp) ? lift[(size_t)m * nfn + (l >> 4) * nfp + ks] : 0.0;
        }
        pack(lift_words, (((size_t)b * (ksl / kg) + kq) * 64 + l) * aw, vals);
      }
}

// {{{ mixed-precision LSERK stage: FP64 state and residual, FP32 RHS in,
// FP32 shadow of the state out

struct LserkMixedArgs {
  OutPtrs q;
  OutPtrs res;
  const float* rhs[kMaxFields];
  float* q32[kMaxFields];
  int64_t n;
  int nf;
  double a, b, dt;
};

__global__ __launch_bounds__(kBlock) void lserk_mixed_kernel(LserkMixedArgs s) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int f = 0; f < s.nf; ++f) {
    double* q = s.q.p[f];
    double* res = s.res.p[f];
    const float* rhs = s.rhs[f];
    float* q32 = s.q32[f];
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < s.n; k += stride) {
      const double r = (s.a != 0.0 ? s.a * res[k] : 0.0) + s.dt * (double)rhs[k];
      res[k] = r;
      const double qn = q[k] + s.b * r;
      q[k] = qn;
      q32[k] = (float)qn;
    }
  }
}

// }}}

// A-operand tiles of the MFMA kernel (see rhs_mfma.h for the layout).  With
// the direct flux the LIFT tiles take its k order: lane group g of k-step ks
// holds face-node pair direct_pair_of(eq, order, ks, g) (zero for a padded
// slot).
void build_tiles(int eq, int order, const double* op, const double* lift, bool direct,
                 double op_scale, std::vector<double>& op_tiles, std::vector<double>& lift_tiles) {
  const int np = np_of(order), nfp = nfp_of(order);
  const int mb = (np + 15) / 16, ks_n = (np + 3) / 4, nfn = kFaces * nfp;
  // A row of lane l in block b: l % 16 for v_mfma_f64_16x16x4f64, l % 4 for
  // the 4x4x4 form a short last block runs on
  auto row = [&](int b, int l) {
    return 16 * b + ((b == mb - 1 && short_last_block(np)) ? (l & 3) : (l & 15));
  };
  op_tiles.assign((size_t)3 * mb * ks_n * 64, 0.0);
  lift_tiles.assign((size_t)mb * nfp * 64, 0.0);
  for (int r = 0; r < 3; ++r)
    for (int b = 0; b < mb; ++b)
      for (int ks = 0; ks < ks_n; ++ks)
        for (int l = 0; l < 64; ++l) {
          const int m = row(b, l), j = (l >> 4) * ks_n + ks;
          if (m < np && j < np)
            op_tiles[(((size_t)r * mb + b) * ks_n + ks) * 64 + l] =
                op_scale * op[((size_t)r * np + m) * np + j];
        }
  if (direct) {
    const int ksd = direct_ksteps_of(eq, order);
    lift_tiles.assign((size_t)mb * ksd * 64, 0.0);
    for (int b = 0; b < mb; ++b)
      for (int ks = 0; ks < ksd; ++ks)
        for (int l = 0; l < 64; ++l) {
          const int m = row(b, l), fp = direct_pair_of(eq, order, ks, l >> 4);
          if (m < np && fp >= 0)
            lift_tiles[((size_t)b * ksd + ks) * 64 + l] = lift[(size_t)m * nfn + fp];
        }
    return;
  }
  for (int b = 0; b < mb; ++b)
    for (int ks = 0; ks < nfp; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int m = row(b, l), fn = (l >> 4) * nfp + ks;
        if (m < np) lift_tiles[((size_t)b * nfp + ks) * 64 + l] = lift[(size_t)m * nfn + fn];
      }
}

// MFMA A tiles of one (Np, Np) matrix in the layout of the D tiles above
// (a short last block on the 4x4x4 form).
std::vector<double> square_tiles(int np, const double* mat) {
  const int mb = (np + 15) / 16, ks_n = (np + 3) / 4;
  std::vector<double> t((size_t)mb * ks_n * 64, 0.0);
  for (int b = 0; b < mb; ++b)
    for (int ks = 0; ks < ks_n; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int m = 16 * b + ((b == mb - 1 && short_last_block(np)) ? (l & 3) : (l & 15));
        const int j = (l >> 4) * ks_n + ks;
        if (m < np && j < np) t[((size_t)b * ks_n + ks) * 64 + l] = mat[(size_t)m * np + j];
      }
  return t;
}

// Tile-ordered copies of the geometry for the MFMA kernels (16-element
// tiles; slots past the last element repeat it):
//   vgeo[t][x][n]        = geo[16 t + n][x]                 (x < 9: rx)
//   fgeo[t][f*16 + n][q] = geo[16 t + n][9 + 4 f + q]       (n, Fscale)
void build_tile_geometry(int64_t k, const double* geo, std::vector<double>& vgeo,
                         std::vector<double>& fgeo) {
  const int64_t ntiles = (k + 15) / 16;
  vgeo.assign((size_t)ntiles * 9 * 16, 0.0);
  fgeo.assign((size_t)ntiles * 64 * 4, 0.0);
  for (int64_t t = 0; t < ntiles; ++t)
    for (int n = 0; n < 16; ++n) {
      const int64_t e = std::min<int64_t>(t * 16 + n, k - 1);
      for (int x = 0; x < 9; ++x) vgeo[((size_t)t * 9 + x) * 16 + n] = geo[e * kGeo + x];
      for (int f = 0; f < kFaces; ++f)
        for (int q = 0; q < 4; ++q)
          fgeo[((size_t)t * 64 + f * 16 + n) * 4 + q] = geo[e * kGeo + 9 + 4 * f + q];
    }
}

// The exterior map in the flux-item order of an MFMA kernel that stages C
// LIFT k-steps per chunk in nchunk chunks:
//   ext_t[t][(ch*C + it)*64 + l] = emap[16 t + p % 16][(p / 16) * Nfp + ks]
// for item = it*64 + l, p = item / C, ks = ch*C + item % C (0 past Nfp).
void build_item_map(int64_t k, int order, int c, int nchunk, const int32_t* emap,
                    std::vector<int32_t>& ext_t) {
  const int nfp = nfp_of(order), nfn = kFaces * nfp;
  const int items = nchunk * c * 64;
  const int64_t ntiles = (k + 15) / 16;
  ext_t.assign((size_t)ntiles * items, 0);
  for (int64_t t = 0; t < ntiles; ++t)
    for (int slot = 0; slot < items; ++slot) {
      const int ch = slot / (c * 64), item = slot % (c * 64);
      const int p = item / c, ks = ch * c + item % c;
      const int64_t e = std::min<int64_t>(t * 16 + (p & 15), k - 1);
      ext_t[(size_t)t * items + slot] = ks < nfp ? emap[e * nfn + (p >> 4) * nfp + ks] : 0;
    }
}

// The exterior map in the direct flux order of the MFMA kernel:
//   ext_d[t][ks][l] = emap[16 t + l % 16][direct_pair_of(eq, order, ks, l / 16)]
// (0, a valid node, for a padded slot; slots past the last element repeat it).
void build_direct_map(int64_t k, int eq, int order, const int32_t* emap,
                      std::vector<int32_t>& ext_d) {
  const int nfn = kFaces * nfp_of(order);
  const int ksd = direct_ksteps_of(eq, order);
  const int64_t ntiles = (k + 15) / 16;
  ext_d.assign((size_t)ntiles * ksd * 64, 0);
  for (int64_t t = 0; t < ntiles; ++t)
    for (int ks = 0; ks < ksd; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int fp = direct_pair_of(eq, order, ks, l >> 4);
        const int64_t e = std::min<int64_t>(t * 16 + (l & 15), k - 1);
        if (fp >= 0) ext_d[((size_t)t * ksd + ks) * 64 + l] = emap[e * nfn + fp];
      }
}

}  // namespace

struct dg_plan {
  int equation = 0;
  int order = 0;
  int dim = 3;
  int np = 0;
  int nfp = 0;
  int nfields = 0;
  int64_t nelem = 0;
  double params[4] = {0, 0, 0, 0};
  int device = 0;
  int grid = 0;          // persistent grid of the RHS kernel
  RhsKernel stage_kernel = nullptr;   // MFMA plans: the LSERK4 stage fused into the RHS
  int stage_grid = 0;
  int lserk_grid = 0;
  KernelInfo kernel;
  // triangles: FP64 and FP32-storage forms of the 2D kernel
  TriKernel<double> tri_kernel = nullptr;
  TriKernel<float> tri_kernel32 = nullptr;
  int variant = 0;
  double* geo = nullptr;
  int32_t* emap = nullptr;
  int32_t* emap_b = nullptr;   // adjoint kernel: in-batch neighbours encoded (kInTile)
  double* dmat_t = nullptr;
  double* lift_t = nullptr;
  int32_t* fmask = nullptr;
  double* op_tiles = nullptr;
  double* lift_tiles = nullptr;
  // MFMA kernels: tile-ordered geometry and exterior maps in flux-item order
  double* vgeo = nullptr;
  double* fgeo = nullptr;
  int32_t* ext_t = nullptr;
  int32_t* ext_d = nullptr;    // direct flux order, where the kernel uses it
  // mixed-precision forms: A tiles and launch grid per precision
  uint32_t* mixed_op[2] = {nullptr, nullptr};
  uint32_t* mixed_lift[2] = {nullptr, nullptr};
  MixedKernel mixed_kernel[2] = {nullptr, nullptr};
  int mixed_grid[2] = {0, 0};
  // tetrahedra: the mixed kernels with the LSERK4 stage fused
  MixedKernel mixed_stage_kernel[2] = {nullptr, nullptr};
  int mixed_stage_grid[2] = {0, 0};
  // locality schedules: 16-element tiles (MFMA and mixed kernels) and
  // batches of kBlock / Np elements (LDS-staged and adjoint kernels)
  int32_t* sched_tile = nullptr;
  int32_t* part_tile = nullptr;
  int32_t* sched_batch = nullptr;
  int32_t* part_batch = nullptr;
  // adjoint (present when the descriptor carried mass_inv and face_mass)
  bool has_vjp = false;
  VjpKernel<double> vjp_kernel = nullptr;
  VjpKernel<float> vjp_kernel32 = nullptr;
  MassKernel<double> mass_kernel = nullptr;
  MassKernel<float> mass_kernel32 = nullptr;
  RhsKernel post_kernel = nullptr;   // tetrahedral wave / Maxwell: fused M post-multiply
  int post_grid = 0;
  double* post_tiles = nullptr;      // A tiles of M_ref
  int vjp_grid = 0;
  int vjp_grid32 = 0;
  int mass_grid = 0;
  double* amat_t = nullptr;      // (d, Np, Np): A_r = D_r^T M at [r][j][i]
  double* face_mass = nullptr;   // (F, Nfp, Nfp)
  double* mass_inv = nullptr;    // (Np, Np)
  double* mass = nullptr;        // (Np, Np) reference mass M
  double* jac = nullptr;         // (K) |J| per element
  void* scratch = nullptr;       // this device's `scratch` (dg_rhs_vjp_noalloc)
  double* ext_fscale = nullptr;  // (K, F)
  int32_t* node_faces = nullptr; // (Np, 3)
  // whether the face geometry satisfies Fscale n = rx^T c_f (checked where a
  // kernel of this plan relies on it)
  bool ref_normals_ok = true;
  // FP64 Euler adjoint on tetrahedra: the MFMA kernel and its A tiles
  VjpMfmaKernel vjp_mfma_kernel = nullptr;
  int vjp_mfma_grid = 0;
  double* vjp_vol_tiles = nullptr;
  double* vjp_face_tiles = nullptr;
  double* vjp_own_tiles = nullptr;
  int32_t* vjp_ext = nullptr;
};

namespace {

void free_plan(dg_plan* p) {
  if (!p) return;
  (void)hipFree(p->geo);
  (void)hipFree(p->emap);
  (void)hipFree(p->emap_b);
  (void)hipFree(p->dmat_t);
  (void)hipFree(p->lift_t);
  (void)hipFree(p->fmask);
  (void)hipFree(p->op_tiles);
  (void)hipFree(p->lift_tiles);
  (void)hipFree(p->vgeo);
  (void)hipFree(p->fgeo);
  (void)hipFree(p->ext_t);
  (void)hipFree(p->ext_d);
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(p->mixed_op[k]);
    (void)hipFree(p->mixed_lift[k]);
  }
  (void)hipFree(p->sched_tile);
  (void)hipFree(p->part_tile);
  (void)hipFree(p->sched_batch);
  (void)hipFree(p->part_batch);
  (void)hipFree(p->amat_t);
  (void)hipFree(p->face_mass);
  (void)hipFree(p->mass_inv);
  (void)hipFree(p->mass);
  (void)hipFree(p->jac);
  (void)hipFree(p->post_tiles);
  (void)hipFree(p->ext_fscale);
  (void)hipFree(p->node_faces);
  (void)hipFree(p->vjp_vol_tiles);
  (void)hipFree(p->vjp_face_tiles);
  (void)hipFree(p->vjp_own_tiles);
  (void)hipFree(p->vjp_ext);
  delete p;
}

template <typename T>
int upload(T** dst, const T* src, size_t count) {
  hipError_t err = hipMalloc((void**)dst, count * sizeof(T));
  if (err != hipSuccess) {
    *dst = nullptr;
    return fail(DG_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(err));
  }
  err = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
  if (err != hipSuccess) return fail(DG_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(err));
  return DG_OK;
}

template <typename T>
void fill_2d_args(const dg_plan* p, Rhs2DArgs<T>& a) {
  a.sched = {p->sched_batch, p->part_batch};
  a.geo = p->geo;
  a.emap = p->emap;
  a.dmat_t = p->dmat_t;
  a.lift_t = p->lift_t;
  a.fmask = p->fmask;
  a.nelem = p->nelem;
  a.p0 = p->params[0];
}

int check_ptrs(const void* const* ptrs, int n, const char* what) {
  if (!ptrs) return fail(DG_ERR_INVALID_ARGUMENT, std::string(what) + " is NULL");
  for (int k = 0; k < n; ++k)
    if (!ptrs[k]) return fail(DG_ERR_INVALID_ARGUMENT, std::string(what) + "[" + std::to_string(k) + "] is NULL");
  return DG_OK;
}

// Whether fields of na bytes at a and nb bytes at b share memory.
bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + nb && y < x + na;
}

// A kernel that gathers neighbour traces of its input fields while it writes
// its output fields needs the outputs to share no memory with the inputs
// (another wave may still read them) or with each other.
int check_disjoint(const void* const* out, const char* out_name, size_t out_bytes,
                   const void* const* in, const char* in_name, size_t in_bytes, int n) {
  for (int c = 0; c < n; ++c)
    for (int c2 = 0; c2 < n; ++c2) {
      if (overlaps(out[c], out_bytes, in[c2], in_bytes))
        return fail(DG_ERR_INVALID_ARGUMENT, std::string(out_name) + " must not overlap " + in_name +
                                                 " (" + out_name + "[" + std::to_string(c) + "], " +
                                                 in_name + "[" + std::to_string(c2) + "])");
      if (c2 > c && overlaps(out[c], out_bytes, out[c2], out_bytes))
        return fail(DG_ERR_INVALID_ARGUMENT, std::string(out_name) + "[" + std::to_string(c) +
                                                 "] and [" + std::to_string(c2) + "] overlap");
    }
  return DG_OK;
}

// Arguments of the tetrahedral FP64 RHS kernels (batch schedule for the
// LDS-staged kernel, tile schedule for the MFMA kernels).
RhsArgs rhs_args(const dg_plan* p, const double* const* in, double* const* out, bool batches) {
  RhsArgs a{};
  for (int c = 0; c < p->nfields; ++c) {
    a.q.p[c] = in[c];
    a.rhs.p[c] = out[c];
  }
  a.sched = batches ? Schedule{p->sched_batch, p->part_batch} : Schedule{p->sched_tile, p->part_tile};
  a.geo = p->geo;
  a.emap = p->emap;
  a.dmat_t = p->dmat_t;
  a.lift_t = p->lift_t;
  a.fmask = p->fmask;
  a.op_tiles = p->op_tiles;
  a.lift_tiles = p->lift_tiles;
  a.vgeo = p->vgeo;
  a.fgeo = p->fgeo;
  a.face_geo_rx = p->ref_normals_ok ? 1 : 0;
  a.ext_t = p->ext_t;
  a.ext_d = p->ext_d;
  a.nelem = p->nelem;
  a.p0 = p->params[0];
  return a;
}

// Arguments of the tetrahedral mixed kernels (precision slot k: 0 FP32, 1 BF16x3).
MixedArgs mixed_args(const dg_plan* p, int k, const float* const* in, float* const* out) {
  MixedArgs a{};
  for (int c = 0; c < p->nfields; ++c) {
    a.q[c] = in[c];
    a.rhs[c] = out[c];
  }
  a.sched = {p->sched_tile, p->part_tile};
  a.geo = p->geo;
  a.emap = p->emap;
  a.fmask = p->fmask;
  a.vgeo = p->vgeo;
  a.fgeo = p->fgeo;
  a.face_geo_rx = p->ref_normals_ok ? 1 : 0;
  a.ext_t = p->ext_t;
  a.ext_d = p->ext_d;
  a.op_tiles = p->mixed_op[k];
  a.lift_tiles = p->mixed_lift[k];
  a.nelem = p->nelem;
  a.p0 = p->params[0];
  return a;
}

}  // namespace

extern "C" {

const char* dg_last_error_string(void) { return g_last_error.c_str(); }

namespace {

// The mixed kernels on the direct flux gather neighbours by buffer loads
// with 32-bit byte offsets: refused where an FP32 field reaches 4 GiB.  The
// wave ones also LIFT the velocity flux in reference normals: refused on a
// geometry that does not allow it.
int check_mixed_geometry(const dg_plan* p, int32_t precision, const char* name) {
  if (p->dim == 2 || !mixed_direct_flux_of(p->equation, precision, p->order)) return DG_OK;
  if (p->nelem * p->np * (int64_t)sizeof(float) >= ((int64_t)1 << 32))
    return fail(DG_ERR_UNSUPPORTED,
                std::string(name) +
                    ": each FP32 field must stay below 4 GiB for this mixed kernel (32-bit "
                    "buffer offsets); use the FP64 form");
  if (p->equation == DG_EQ_WAVE && !p->ref_normals_ok)
    return fail(DG_ERR_UNSUPPORTED,
                std::string(name) +
                    ": the mixed wave kernel LIFTs its velocity flux in reference normals, and "
                    "this plan's face geometry is not the affine image of the reference "
                    "element's (Fscale n != rx^T c_f); use the FP64 form");
  return DG_OK;
}

}  // namespace

int dg_rhs_mixed(const dg_plan* p, int32_t precision, const float* const* in, float* const* out,
                 void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (precision != DG_PREC_FP32 && precision != DG_PREC_BF16X3)
    return fail(DG_ERR_UNSUPPORTED, "mixed precision " + std::to_string(precision) +
                                        " not supported (FP32 or BF16X3)");
  int rc;
  if ((rc = check_mixed_geometry(p, precision, "dg_rhs_mixed"))) return rc;
  if ((rc = check_ptrs((const void* const*)in, p->nfields, "fields_in")) ||
      (rc = check_ptrs((const void* const*)out, p->nfields, "fields_out")) ||
      (rc = check_disjoint((const void* const*)out, "fields_out",
                           (size_t)(p->nelem * p->np) * sizeof(float), (const void* const*)in,
                           "fields_in", (size_t)(p->nelem * p->np) * sizeof(float), p->nfields)))
    return rc;
  if (p->dim == 2) {
    // triangles: FP32 state, every operation in FP64 (both precisions)
    Rhs2DArgs<float> a{};
    for (int c = 0; c < p->nfields; ++c) {
      a.q[c] = in[c];
      a.rhs[c] = out[c];
    }
    fill_2d_args(p, a);
    hipLaunchKernelGGL(p->tri_kernel32, dim3(p->grid), dim3(kBlock), 0, (hipStream_t)stream, a);
    DG_HIP_CHECK(hipGetLastError());
    return DG_OK;
  }
  const int k = precision == DG_PREC_FP32 ? 0 : 1;
  const MixedArgs a = mixed_args(p, k, in, out);
  hipLaunchKernelGGL(p->mixed_kernel[k], dim3(p->mixed_grid[k]), dim3(kBlock), 0,
                     (hipStream_t)stream, a);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}

int dg_rhs_lserk_stage_mixed(const dg_plan* p, int32_t precision, double a, double b, double dt,
                             const float* const* q32_in, float* const* q32_out, double* const* q,
                             double* const* res, void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (precision != DG_PREC_FP32 && precision != DG_PREC_BF16X3)
    return fail(DG_ERR_UNSUPPORTED, "mixed precision " + std::to_string(precision) +
                                        " not supported (FP32 or BF16X3)");
  if (p->dim == 2)
    return fail(DG_ERR_UNSUPPORTED,
                "dg_rhs_lserk_stage_mixed: tetrahedral plans only (use dg_rhs_mixed and "
                "dg_lserk_stage_mixed)");
  int rc;
  if ((rc = check_mixed_geometry(p, precision, "dg_rhs_lserk_stage_mixed"))) return rc;
  if ((rc = check_ptrs((const void* const*)q32_in, p->nfields, "q32_in")) ||
      (rc = check_ptrs((const void* const*)q32_out, p->nfields, "q32_out")) ||
      (rc = check_ptrs((const void* const*)q, p->nfields, "q")) ||
      (rc = check_ptrs((const void* const*)res, p->nfields, "res")))
    return rc;
  const size_t n = (size_t)(p->nelem * p->np);
  const size_t b4 = n * sizeof(float), b8 = n * sizeof(double);
  if ((rc = check_disjoint((const void* const*)q32_out, "q32_out", b4,
                           (const void* const*)q32_in, "q32_in", b4, p->nfields)) ||
      (rc = check_disjoint((const void* const*)q, "q", b8, (const void* const*)q32_in, "q32_in", b4,
                           p->nfields)) ||
      (rc = check_disjoint((const void* const*)res, "res", b8, (const void* const*)q32_in,
                           "q32_in", b4, p->nfields)) ||
      (rc = check_disjoint((const void* const*)res, "res", b8, (const void* const*)q, "q", b8,
                           p->nfields)) ||
      (rc = check_disjoint((const void* const*)q32_out, "q32_out", b4, (const void* const*)q, "q",
                           b8, p->nfields)) ||
      (rc = check_disjoint((const void* const*)q32_out, "q32_out", b4, (const void* const*)res,
                           "res", b8, p->nfields)))
    return rc;
  const int k = precision == DG_PREC_FP32 ? 0 : 1;
  float* none[kMaxFields] = {};
  MixedArgs m = mixed_args(p, k, q32_in, none);
  for (int c = 0; c < p->nfields; ++c) {
    m.st_q[c] = q[c];
    m.st_res[c] = res[c];
    m.st_q32[c] = q32_out[c];
  }
  m.st_a = a;
  m.st_b = b;
  m.st_dt = dt;
  hipLaunchKernelGGL(p->mixed_stage_kernel[k], dim3(p->mixed_stage_grid[k]), dim3(kBlock), 0,
                     (hipStream_t)stream, m);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}

int dg_lserk_stage_mixed(const dg_plan* p, double a, double b, double dt, double* const* q,
                         double* const* res, const float* const* rhs, float* const* q32,
                         void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  int rc;
  if ((rc = check_ptrs((const void* const*)q, p->nfields, "q")) ||
      (rc = check_ptrs((const void* const*)res, p->nfields, "res")) ||
      (rc = check_ptrs((const void* const*)rhs, p->nfields, "rhs")) ||
      (rc = check_ptrs((const void* const*)q32, p->nfields, "q32")))
    return rc;
  LserkMixedArgs s{};
  for (int c = 0; c < p->nfields; ++c) {
    s.q.p[c] = q[c];
    s.res.p[c] = res[c];
    s.rhs[c] = rhs[c];
    s.q32[c] = q32[c];
  }
  s.n = p->nelem * p->np;
  s.nf = p->nfields;
  s.a = a;
  s.b = b;
  s.dt = dt;
  hipLaunchKernelGGL(lserk_mixed_kernel, dim3(p->lserk_grid), dim3(kBlock), 0,
                     (hipStream_t)stream, s);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}

int dg_plan_create(const dg_plan_desc* d, dg_plan** out) {
  g_last_error.clear();
  if (!d || !out) return fail(DG_ERR_INVALID_ARGUMENT, "NULL descriptor or output");
  *out = nullptr;
  if (d->variant < DG_KERNEL_AUTO || d->variant > DG_KERNEL_LDS)
    return fail(DG_ERR_INVALID_ARGUMENT, "unknown kernel variant " + std::to_string(d->variant));
  if (d->equation < DG_EQ_WAVE || d->equation > DG_EQ_MAXWELL)
    return fail(DG_ERR_UNSUPPORTED, "unknown equation id " + std::to_string(d->equation));
  if (d->order < 1 || d->order > 4)
    return fail(DG_ERR_UNSUPPORTED, "order " + std::to_string(d->order) + " not supported (1-4)");
  if (d->nelements < 1) return fail(DG_ERR_INVALID_ARGUMENT, "nelements must be >= 1");
  if (d->dim != 0 && d->dim != 2 && d->dim != 3)
    return fail(DG_ERR_UNSUPPORTED, "dim " + std::to_string(d->dim) + " not supported (2 or 3)");
  const bool tri = d->dim == 2;
  const int np = tri ? np2_of(d->order) : np_of(d->order);
  const int nfp = tri ? nfp2_of(d->order) : nfp_of(d->order);
  const int nfaces = tri ? kFaces2 : kFaces;
  const int ndir = tri ? 2 : 3;
  const int geo_stride = tri ? kGeo2 : kGeo;
  if ((int64_t)d->nelements * np >= (int64_t)INT32_MAX ||
      (int64_t)d->nelements * nfaces * nfp >= (int64_t)INT32_MAX)
    return fail(DG_ERR_UNSUPPORTED, "mesh too large for 32-bit node indices");
  if (!d->d_weak || !d->d_strong || !d->lift || !d->face_nodes || !d->geometry || !d->ext_map)
    return fail(DG_ERR_INVALID_ARGUMENT, "NULL operator or mesh array in descriptor");
  if (tri && d->variant == DG_KERNEL_MFMA)
    return fail(DG_ERR_UNSUPPORTED, "MFMA kernel: tetrahedral plans only (dim 2 runs the "
                                    "LDS-staged kernel)");

  const int64_t k = d->nelements;
  const int nfn = nfaces * nfp;
  // validate the exterior map on the host: every index in range
  for (int64_t t = 0; t < k * nfn; ++t) {
    const int64_t v = d->ext_map[t];
    const int64_t idx = v >= 0 ? v : -v - 1;
    if (idx >= k * np)
      return fail(DG_ERR_INVALID_ARGUMENT, "ext_map entry out of range at " + std::to_string(t));
  }
  for (int t = 0; t < nfn; ++t)
    if (d->face_nodes[t] < 0 || d->face_nodes[t] >= np)
      return fail(DG_ERR_INVALID_ARGUMENT, "face_nodes entry out of range");
  // adjoint data, built and checked before anything touches the device
  const bool want_vjp = d->mass_inv || d->face_mass;
  std::vector<double> fm, xfs, amat_t, mass, jac;
  std::vector<int32_t> nf;
  if (want_vjp) {
    if (!d->mass_inv || !d->face_mass)
      return fail(DG_ERR_INVALID_ARGUMENT, "mass_inv and face_mass must be given together");
    const std::string msg =
        tri ? build_adjoint_data(d, np, nfp, kFaces2, kGeo2, 4 + 2, 3, fm, nf, xfs)
            : build_adjoint_data(d, np, nfp, kFaces, kGeo, 9 + 3, 4, fm, nf, xfs);
    if (!msg.empty()) return fail(DG_ERR_INVALID_ARGUMENT, msg);
    // A_r = D_r^T M, stored [r][j][i] = A_r[i][j]: the adjoint's volume
    // operator acting on mu = M^{-1} lam
    if (!invert(np, d->mass_inv, mass)) return fail(DG_ERR_INVALID_ARGUMENT, "mass_inv is singular");
    // |J| = 1 / |det(dr/dx)| per element, from the rx block of the geometry
    // rows (the element mass matrix is |J| M)
    jac.resize((size_t)k);
    for (int64_t e = 0; e < k; ++e) {
      const double* g = d->geometry + e * (tri ? kGeo2 : kGeo);
      const double det = tri ? g[0] * g[3] - g[1] * g[2]
                             : g[0] * (g[4] * g[8] - g[5] * g[7]) -
                                   g[1] * (g[3] * g[8] - g[5] * g[6]) +
                                   g[2] * (g[3] * g[7] - g[4] * g[6]);
      if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
        return fail(DG_ERR_INVALID_ARGUMENT, "geometry: singular rx block");
      jac[(size_t)e] = 1.0 / std::fabs(det);
    }
    const double* dop = d->equation == DG_EQ_MAXWELL ? d->d_strong : d->d_weak;
    amat_t.assign((size_t)ndir * np * np, 0.0);
    for (int r = 0; r < ndir; ++r)
      for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
          double acc = 0.0;
          for (int kk = 0; kk < np; ++kk)
            acc += dop[((size_t)r * np + kk) * np + i] * mass[(size_t)kk * np + j];
          amat_t[((size_t)r * np + j) * np + i] = acc;
        }
  }

  dg_plan* p = new (std::nothrow) dg_plan;
  if (!p) return fail(DG_ERR_OUT_OF_MEMORY, "host allocation failed");
  p->equation = d->equation;
  p->order = d->order;
  p->dim = tri ? 2 : 3;
  p->np = np;
  p->nfp = nfp;
  p->nfields = tri ? nfields2_of(d->equation) : nfields_of(d->equation);
  p->nelem = k;
  for (int t = 0; t < 4; ++t) p->params[t] = d->params[t];
  if (tri) {
    p->variant = DG_KERNEL_LDS;
    p->tri_kernel = pick_2d<double>(d->equation, d->order);
    p->tri_kernel32 = pick_2d<float>(d->equation, d->order);
    p->kernel = {nullptr, kBlock, kBlock / np};
  } else {
    p->variant = d->variant == DG_KERNEL_AUTO ? auto_variant(d->equation, d->order) : d->variant;
    if (p->variant == DG_KERNEL_MFMA && intile_neighbours_of(d->equation, d->order) &&
        d->nelements * np * 8 >= ((int64_t)1 << 32)) {
      if (d->variant == DG_KERNEL_MFMA) {
        free_plan(p);
        return fail(DG_ERR_UNSUPPORTED,
                    "MFMA kernel: each state field must stay below 4 GiB at this order (32-bit "
                    "buffer offsets); the LDS-staged kernel takes larger meshes");
      }
      p->variant = DG_KERNEL_LDS;
    }
    // the kernels that LIFT the wave operator's velocity flux in reference
    // normals (the FP64 MFMA kernel; the mixed kernels where they run the
    // direct flux) need Fscale n = rx^T c_f on every face; the MFMA kernel
    // of the other operators forms its face geometry from rx where it holds
    p->ref_normals_ok = geometry_fits_ref_normals(d->nelements, d->geometry);
    if (p->variant == DG_KERNEL_MFMA && ref_normal_lift_of(d->equation, d->order) &&
        !p->ref_normals_ok) {
      if (d->variant == DG_KERNEL_MFMA) {
        free_plan(p);
        return fail(DG_ERR_INVALID_ARGUMENT,
                    "MFMA kernel: the face geometry is not the affine image of the reference "
                    "element's (Fscale n != rx^T c_f); the LDS-staged kernel takes any geometry");
      }
      p->variant = DG_KERNEL_LDS;
    }
    p->kernel = p->variant == DG_KERNEL_LDS ? pick_lds(d->equation, d->order)
                                            : pick_mfma(d->equation, d->order);
  }

  int rc;
  if (hipGetDevice(&p->device) != hipSuccess) {
    free_plan(p);
    return fail(DG_ERR_HIP, "hipGetDevice failed");
  }
  // transposed operators: dmat_t[r][j][i] = D[r][i][j]; lift_t[fn][i] = LIFT[i][fn]
  const double* dsrc = d->equation == DG_EQ_MAXWELL ? d->d_strong : d->d_weak;
  std::vector<double> dt((size_t)ndir * np * np), lt((size_t)nfn * np);
  for (int r = 0; r < ndir; ++r)
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) dt[(r * np + j) * np + i] = dsrc[(r * np + i) * np + j];
  for (int i = 0; i < np; ++i)
    for (int fn = 0; fn < nfn; ++fn) lt[fn * np + i] = d->lift[i * nfn + fn];
  if (!tri) {
    std::vector<double> op_tiles, lift_tiles;
    // the wave kernel's D tiles carry the operator's factor -c
    build_tiles(d->equation, d->order, dsrc, d->lift, direct_flux_of(d->equation, d->order),
                d->equation == DG_EQ_WAVE ? -d->params[0] : 1.0, op_tiles, lift_tiles);
    if ((rc = upload(&p->op_tiles, op_tiles.data(), op_tiles.size())) ||
        (rc = upload(&p->lift_tiles, lift_tiles.data(), lift_tiles.size()))) {
      free_plan(p);
      return rc;
    }
  }
  if (!tri) {
    // tile-ordered geometry and flux-item-ordered exterior map of the FP64
    // MFMA kernel, which the mixed kernels read too (the same data as
    // geometry / ext_map, laid out so that each of the kernels' loads is
    // contiguous; measured on MI355X: FP64 wave P3 -10%, Maxwell P4 -6%,
    // identical results)
    std::vector<double> vgeo, fgeo;
    std::vector<int32_t> ext_t;
    build_tile_geometry(k, d->geometry, vgeo, fgeo);
    const int c = flux_chunk_of(d->equation, d->order);
    build_item_map(k, d->order, c, (nfp + c - 1) / c, d->ext_map, ext_t);
    if ((rc = upload(&p->vgeo, vgeo.data(), vgeo.size())) ||
        (rc = upload(&p->fgeo, fgeo.data(), fgeo.size())) ||
        (rc = upload(&p->ext_t, ext_t.data(), ext_t.size()))) {
      free_plan(p);
      return rc;
    }
    // the direct map: for the FP64 kernel where it runs the direct flux, and
    // for the FP32 mixed kernel alone where only that one does (wave P1,
    // Euler P1-P3, Maxwell P1, P3), with the in-tile entries, since that
    // kernel stages its tile
    const bool mixed_only = !direct_flux_of(d->equation, d->order) &&
                            mixed_direct_flux_of(d->equation, DG_PREC_FP32, d->order);
    if (direct_flux_of(d->equation, d->order) || mixed_only) {
      std::vector<int32_t> ext_d;
      build_direct_map(k, d->equation, d->order, d->ext_map, ext_d);
      if (intile_neighbours_of(d->equation, d->order) || mixed_only) {
        // neighbours inside the tile point into the staged state
        const int ksd = direct_ksteps_of(d->equation, d->order);
        const int64_t nt = (k + 15) / 16;
        for (int64_t t = 0; t < nt; ++t)
          for (int ks = 0; ks < ksd; ++ks)
            for (int l = 0; l < 64; ++l) {
              int32_t& v = ext_d[((size_t)t * ksd + ks) * 64 + l];
              const int64_t e = v / np;
              if (v >= 0 && e / 16 == t && direct_pair_of(d->equation, d->order, ks, l >> 4) >= 0)
                v = kInTile | (int32_t)((e % 16) << 8) | (v % np);
            }
      }
      if ((rc = upload(&p->ext_d, ext_d.data(), ext_d.size()))) {
        free_plan(p);
        return rc;
      }
    }
  }
  if ((rc = upload(&p->dmat_t, dt.data(), dt.size())) ||
      (rc = upload(&p->lift_t, lt.data(), lt.size())) ||
      (rc = upload(&p->fmask, d->face_nodes, (size_t)nfn)) ||
      (rc = upload(&p->geo, d->geometry, (size_t)k * geo_stride)) ||
      (rc = upload(&p->emap, d->ext_map, (size_t)k * nfn))) {
    free_plan(p);
    return rc;
  }
  {
    std::vector<int32_t> order;
    std::array<int32_t, 9> part;
    build_schedule(k, 16, d->centroids, order, part);
    if ((rc = upload(&p->sched_tile, order.data(), order.size())) ||
        (rc = upload(&p->part_tile, part.data(), part.size()))) {
      free_plan(p);
      return rc;
    }
    build_schedule(k, kBlock / np, d->centroids, order, part);
    if ((rc = upload(&p->sched_batch, order.data(), order.size())) ||
        (rc = upload(&p->part_batch, part.data(), part.size()))) {
      free_plan(p);
      return rc;
    }
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) {
    free_plan(p);
    return fail(DG_ERR_HIP, "hipGetDeviceProperties failed");
  }
  int per_cu = 0;
  const void* main_fn = tri ? (const void*)p->tri_kernel : (const void*)p->kernel.fn;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, main_fn, p->kernel.threads, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int64_t nbatch = (k + p->kernel.batch_elems - 1) / p->kernel.batch_elems;
  int64_t grid = (int64_t)prop.multiProcessorCount * per_cu;
  if (grid > nbatch) grid = nbatch;
  // the kernel's XCD remap needs a multiple of 8 blocks; spare blocks idle
  grid = (grid + 7) / 8 * 8;
  p->grid = (int)grid;
  p->lserk_grid = prop.multiProcessorCount * 8;
  if (!tri && p->variant == DG_KERNEL_MFMA) {
    p->stage_kernel = pick_mfma_stage(d->equation, d->order);
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)p->stage_kernel,
                                                      p->kernel.threads, 0) != hipSuccess ||
        occ < 1)
      occ = 1;
    int64_t sg = (int64_t)prop.multiProcessorCount * occ;
    if (sg > nbatch) sg = nbatch;
    p->stage_grid = (int)((sg + 7) / 8 * 8);
  }
  for (int k = 0; k < 2 && !tri; ++k) {
    const int prec = k == 0 ? DG_PREC_FP32 : DG_PREC_BF16X3;
    std::vector<uint32_t> ow, lw;
    build_mixed_tiles(prec, bf16_kgroup_of(d->equation, d->order), d->equation, d->order, dsrc,
                      d->lift, ow, lw);
    if ((rc = upload(&p->mixed_op[k], ow.data(), ow.size())) ||
        (rc = upload(&p->mixed_lift[k], lw.data(), lw.size()))) {
      free_plan(p);
      return rc;
    }
    p->mixed_kernel[k] = pick_mixed(prec, d->equation, d->order);
    p->mixed_stage_kernel[k] = pick_mixed(prec, d->equation, d->order, true);
    for (int st = 0; st < 2; ++st) {
      const MixedKernel fn = st ? p->mixed_stage_kernel[k] : p->mixed_kernel[k];
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)fn, kBlock, 0) !=
              hipSuccess ||
          occ < 1)
        occ = 1;
      int64_t mg = (int64_t)prop.multiProcessorCount * occ;
      const int64_t mbatch = (((d->nelements + 15) / 16) + 3) / 4;
      if (mg > mbatch) mg = mbatch;
      (st ? p->mixed_stage_grid[k] : p->mixed_grid[k]) = (int)((mg + 7) / 8 * 8);
    }
  }
  if (want_vjp) {
    if ((rc = upload(&p->amat_t, amat_t.data(), amat_t.size())) ||
        (rc = upload(&p->face_mass, fm.data(), fm.size())) ||
        (rc = upload(&p->mass_inv, d->mass_inv, (size_t)np * np)) ||
        (rc = upload(&p->mass, mass.data(), mass.size())) ||
        (rc = upload(&p->jac, jac.data(), jac.size())) ||
        (rc = upload(&p->ext_fscale, xfs.data(), xfs.size())) ||
        (rc = upload(&p->node_faces, nf.data(), nf.size()))) {
      free_plan(p);
      return rc;
    }
    if (!tri && d->equation != DG_EQ_EULER && p->variant == DG_KERNEL_MFMA) {
      const std::vector<double> mt = square_tiles(np, mass.data());
      if ((rc = upload(&p->post_tiles, mt.data(), mt.size()))) {
        free_plan(p);
        return rc;
      }
      p->post_kernel = pick_mfma_post(d->equation, d->order);
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)p->post_kernel,
                                                        p->kernel.threads, 0) != hipSuccess ||
          occ < 1)
        occ = 1;
      int64_t pg = (int64_t)prop.multiProcessorCount * occ;
      const int64_t tbatch = (k + p->kernel.batch_elems - 1) / p->kernel.batch_elems;
      if (pg > tbatch) pg = tbatch;
      p->post_grid = (int)((pg + 7) / 8 * 8);
    }
    p->vjp_kernel = pick_vjp<double>(d->equation, d->order, p->dim);
    p->vjp_kernel32 = pick_vjp<float>(d->equation, d->order, p->dim);
    {
      // the adjoint kernel's exterior map: neighbours inside the element's
      // batch of kBlock / Np elements point into the batch's staged state
      const int eb = kBlock / np;
      std::vector<int32_t> mb(d->ext_map, d->ext_map + (size_t)k * nfn);
      for (int64_t e = 0; e < k; ++e)
        for (int t = 0; t < nfn; ++t) {
          int32_t& v = mb[(size_t)e * nfn + t];
          if (v >= 0 && v / np / eb == e / eb)
            v = kInTile | (int32_t)(((v / np) % eb) << 8) | (v % np);
        }
      if ((rc = upload(&p->emap_b, mb.data(), mb.size()))) {
        free_plan(p);
        return rc;
      }
    }
    if (!tri && d->equation == DG_EQ_EULER) {
      std::vector<double> vt, ft, ot;
      std::vector<int32_t> xv;
      build_vjp_mfma_tables(k, np, nfp, amat_t.data(), fm.data(), d->face_nodes, d->ext_map,
                            vjp_mfma_intile(d->order), vt, ft, ot, xv);
      if ((rc = upload(&p->vjp_vol_tiles, vt.data(), vt.size())) ||
          (rc = upload(&p->vjp_face_tiles, ft.data(), ft.size())) ||
          (rc = upload(&p->vjp_own_tiles, ot.data(), ot.size())) ||
          (rc = upload(&p->vjp_ext, xv.data(), xv.size()))) {
        free_plan(p);
        return rc;
      }
      p->vjp_mfma_kernel = pick_vjp_mfma(d->order);
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)p->vjp_mfma_kernel, 256,
                                                        0) != hipSuccess ||
          occ < 1)
        occ = 1;
      int64_t vg = (int64_t)prop.multiProcessorCount * occ;
      const int64_t tbatch = ((k + 15) / 16 + 3) / 4;
      if (vg > tbatch) vg = tbatch;
      p->vjp_mfma_grid = (int)((vg + 7) / 8 * 8);
    }
    p->mass_kernel = pick_mass<double>(np, p->nfields);
    p->mass_kernel32 = pick_mass<float>(np, p->nfields);
    const int64_t vbatch = (k + kBlock / np - 1) / (kBlock / np);
    auto grid_of = [&](const void* fn) {
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kBlock, 0) != hipSuccess || occ < 1)
        occ = 1;
      int64_t g = (int64_t)prop.multiProcessorCount * occ;
      if (g > vbatch) g = vbatch;
      return (int)((g + 7) / 8 * 8);
    };
    if (p->vjp_kernel) p->vjp_grid = grid_of((const void*)p->vjp_kernel);
    p->vjp_grid32 = grid_of((const void*)p->vjp_kernel32);
    {
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)p->mass_kernel,
                                                        kBlock, 0) != hipSuccess ||
          occ < 1)
        occ = 1;
      int64_t mg = (int64_t)prop.multiProcessorCount * occ;
      p->mass_grid = (int)(mg < vbatch ? mg : vbatch);
    }
    // set-up of dg_rhs_vjp_noalloc: this device's address of `scratch`
    if ((rc = scratch_address(p->device, &p->scratch)) != DG_OK) {
      free_plan(p);
      return rc;
    }
    p->has_vjp = true;
  }
  *out = p;
  return DG_OK;
}

int dg_plan_destroy(dg_plan* p) {
  free_plan(p);
  return DG_OK;
}

int dg_plan_info(const dg_plan* p, int32_t* nfields, int32_t* np, int32_t* nfp, int64_t* nelem) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (nfields) *nfields = p->nfields;
  if (np) *np = p->np;
  if (nfp) *nfp = p->nfp;
  if (nelem) *nelem = p->nelem;
  return DG_OK;
}

int dg_plan_variant(const dg_plan* p, int32_t* variant) {
  if (!p || !variant) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan or output");
  *variant = p->variant;
  return DG_OK;
}

int dg_rhs(const dg_plan* p, int32_t precision, const double* const* in, double* const* out,
           void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (precision != DG_PREC_FP64)
    return fail(DG_ERR_UNSUPPORTED, "precision " + std::to_string(precision) + " not supported");
  int rc;
  if ((rc = check_ptrs((const void* const*)in, p->nfields, "fields_in")) ||
      (rc = check_ptrs((const void* const*)out, p->nfields, "fields_out")) ||
      (rc = check_disjoint((const void* const*)out, "fields_out",
                           (size_t)(p->nelem * p->np) * sizeof(double), (const void* const*)in,
                           "fields_in", (size_t)(p->nelem * p->np) * sizeof(double), p->nfields)))
    return rc;
  if (p->dim == 2) {
    Rhs2DArgs<double> a{};
    for (int c = 0; c < p->nfields; ++c) {
      a.q[c] = in[c];
      a.rhs[c] = out[c];
    }
    fill_2d_args(p, a);
    hipLaunchKernelGGL(p->tri_kernel, dim3(p->grid), dim3(kBlock), 0, (hipStream_t)stream, a);
    DG_HIP_CHECK(hipGetLastError());
    return DG_OK;
  }
  const RhsArgs a = rhs_args(p, in, out, p->variant == DG_KERNEL_LDS);
  hipLaunchKernelGGL(p->kernel.fn, dim3(p->grid), dim3(p->kernel.threads), 0,
                     (hipStream_t)stream, a);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}

// The MFMA kernel with the stage fused: res = alpha res + dt L(in);
// out = beta in + b res.  names: in, out, res, entry point.
static int fused_stage_impl(const dg_plan* p, double alpha, double beta, double b, double dt,
                            const double* const* in, double* const* out, double* const* res,
                            void* stream, const char* const names[4]) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (!p->stage_kernel)
    return fail(DG_ERR_UNSUPPORTED, std::string(names[3]) +
                                        ": tetrahedral plans on the MFMA kernel only (use dg_rhs "
                                        "and the stage kernel)");
  int rc;
  if ((rc = check_ptrs((const void* const*)in, p->nfields, names[0])) ||
      (rc = check_ptrs((const void* const*)out, p->nfields, names[1])) ||
      (rc = check_ptrs((const void* const*)res, p->nfields, names[2])))
    return rc;
  const size_t bytes = (size_t)(p->nelem * p->np) * sizeof(double);
  if ((rc = check_disjoint((const void* const*)out, names[1], bytes, (const void* const*)in,
                           names[0], bytes, p->nfields)) ||
      (rc = check_disjoint((const void* const*)res, names[2], bytes, (const void* const*)in,
                           names[0], bytes, p->nfields)) ||
      (rc = check_disjoint((const void* const*)res, names[2], bytes, (const void* const*)out,
                           names[1], bytes, p->nfields)))
    return rc;
  double* none[kMaxFields] = {};
  RhsArgs r = rhs_args(p, in, none, false);
  for (int c = 0; c < p->nfields; ++c) {
    r.stage_res.p[c] = res[c];
    r.stage_qout.p[c] = out[c];
  }
  r.stage_a = alpha;
  r.stage_b = b;
  r.stage_dt = dt;
  r.stage_beta = beta;
  hipLaunchKernelGGL(p->stage_kernel, dim3(p->stage_grid), dim3(p->kernel.threads), 0,
                     (hipStream_t)stream, r);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}

int dg_rhs_lserk_stage(const dg_plan* p, double a, double b, double dt, const double* const* q_in,
                       double* const* q_out, double* const* res, void* stream) {
  static const char* const names[4] = {"q_in", "q_out", "res", "dg_rhs_lserk_stage"};
  return fused_stage_impl(p, a, 1.0, b, dt, q_in, q_out, res, stream, names);
}

int dg_rhs_lserk_stage_adjoint(const dg_plan* p, double a, double b, double dt,
                               const double* const* s_in, double* const* s_out,
                               double* const* q_hat, void* stream) {
  static const char* const names[4] = {"s_in", "s_out", "q_hat", "dg_rhs_lserk_stage_adjoint"};
  if (p && p->equation == DG_EQ_EULER)
    return fail(DG_ERR_UNSUPPORTED, "dg_rhs_lserk_stage_adjoint: the Euler operator is not linear");
  return fused_stage_impl(p, 1.0, a, b, dt, s_in, s_out, q_hat, stream, names);
}

int dg_lserk_stage(const dg_plan* p, double a, double b, double dt, double* const* q,
                   double* const* res, const double* const* rhs, void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  int rc;
  if ((rc = check_ptrs((const void* const*)q, p->nfields, "q")) ||
      (rc = check_ptrs((const void* const*)res, p->nfields, "res")) ||
      (rc = check_ptrs((const void* const*)rhs, p->nfields, "rhs")))
    return rc;
  LserkArgs s{};
  for (int c = 0; c < p->nfields; ++c) {
    s.q.p[c] = q[c];
    s.res.p[c] = res[c];
    s.rhs.p[c] = rhs[c];
  }
  s.n = p->nelem * p->np;
  s.nf = p->nfields;
  s.a = a;
  s.b = b;
  s.dt = dt;
  hipLaunchKernelGGL(lserk_stage_kernel, dim3(p->lserk_grid), dim3(kBlock), 0,
                     (hipStream_t)stream, s);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}


}  // extern "C"

namespace {

// Signs of the field reflection R of the linear operators' adjoint: +1 on u
// (wave), E (Maxwell) and E_z (TM), -1 on the velocity and H components.
void adjoint_reflection(const dg_plan* p, double* sign) {
  for (int c = 0; c < p->nfields; ++c) {
    const bool plus = (p->equation == DG_EQ_MAXWELL && p->dim != 2) ? c < 3 : c == 0;
    sign[c] = plus ? 1.0 : -1.0;
  }
}

// Vector-Jacobian product.
//
// FP64 wave and Maxwell on tetrahedra: with exact quadrature on affine
// elements the weak form B = M L of these operators is a skew-symmetric
// central part plus a symmetric upwind penalty, and the reflection R (the
// sign flip of the velocity / magnetic field) negates exactly the central
// part, so B^T = R B R and
//     L^T lam = R M L R M^{-1} lam
// (M the global mass matrix, |J_k| M_ref per element).  The scaled inverse
// mass goes into the workspace; then, for an MFMA plan, one launch of the MFMA
// kernel with R |J| M_ref fused into its output phase writes q_bar, and for an
// LDS-staged plan the forward kernel writes q_bar and a mass pass scales it in
// place.  Measured on MI355X, against the dedicated adjoint kernel: wave P3
// 259 -> 156 us, wave P4 366 -> 220 us, Maxwell P4 711 -> 439 us; Maxwell P2
// (162k tets) 122 us on the MFMA kernel, 158 us on the LDS-staged one.
//
// Euler, triangles and the FP32 adjoint: mu = M_ref^{-1} lam into the
// workspace, then the dedicated adjoint kernel (on triangles the two mass
// passes cost more than the adjoint kernel: wave P1 177 against 159 us).
template <typename T>
int rhs_vjp_impl(const dg_plan* p, const T* const* q, const T* const* rhs_bar, T* const* q_bar,
                 T* const* workspace, void* stream, const char* name) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (!p->has_vjp)
    return fail(DG_ERR_UNSUPPORTED, std::string(name) +
                                        ": plan created without mass_inv / face_mass: no adjoint");
  int rc;
  if (p->equation == DG_EQ_EULER && (rc = check_ptrs((const void* const*)q, p->nfields, "q")))
    return rc;
  if ((rc = check_ptrs((const void* const*)rhs_bar, p->nfields, "rhs_bar")) ||
      (rc = check_ptrs((const void* const*)q_bar, p->nfields, "q_bar")) ||
      (rc = check_ptrs((const void* const*)workspace, p->nfields, "workspace")))
    return rc;
  // The second launch gathers neighbour values of the workspace (and, for
  // Euler, of q) while it writes q_bar, and the first writes the workspace
  // before q is read.  The mass pass into the workspace is safe in place on
  // the same field, so workspace[c] may be rhs_bar[c] but overlap nothing
  // else of rhs_bar.
  const size_t fbytes = (size_t)(p->nelem * p->np) * sizeof(T);
  if ((rc = check_disjoint((const void* const*)q_bar, "q_bar", fbytes,
                           (const void* const*)workspace, "workspace", fbytes, p->nfields)))
    return rc;
  if (p->equation == DG_EQ_EULER &&
      ((rc = check_disjoint((const void* const*)q_bar, "q_bar", fbytes, (const void* const*)q, "q",
                            fbytes, p->nfields)) ||
       (rc = check_disjoint((const void* const*)workspace, "workspace", fbytes,
                            (const void* const*)q, "q", fbytes, p->nfields))))
    return rc;
  for (int c = 0; c < p->nfields; ++c)
    for (int c2 = 0; c2 < p->nfields; ++c2) {
      if (!(c == c2 && (const void*)workspace[c] == (const void*)rhs_bar[c]) &&
          overlaps(workspace[c], fbytes, rhs_bar[c2], fbytes))
        return fail(DG_ERR_INVALID_ARGUMENT,
                    std::string(name) + ": workspace[" + std::to_string(c) +
                        "] overlaps rhs_bar[" + std::to_string(c2) + "]");
      if (c2 > c && overlaps(workspace[c], fbytes, workspace[c2], fbytes))
        return fail(DG_ERR_INVALID_ARGUMENT,
                    std::string(name) + ": workspace[" + std::to_string(c) + "] and [" +
                        std::to_string(c2) + "] overlap");
    }
  constexpr bool f32 = sizeof(T) == 4;
  MassArgs<T> m{};
  for (int c = 0; c < p->nfields; ++c) {
    m.in[c] = rhs_bar[c];
    m.out[c] = workspace[c];
    m.sign[c] = 1.0;
  }
  m.matrix = p->mass_inv;
  m.nelem = p->nelem;
  MassKernel<T> mk;
  if constexpr (f32)
    mk = p->mass_kernel32;
  else
    mk = p->mass_kernel;

  if constexpr (!f32) {
    if (p->equation != DG_EQ_EULER && p->dim != 2) {
      m.jac = p->jac;
      m.jac_power = -1;
      adjoint_reflection(p, m.sign);
      hipLaunchKernelGGL(mk, dim3(p->mass_grid), dim3(kBlock), 0, (hipStream_t)stream, m);
      DG_HIP_CHECK(hipGetLastError());
      if (p->variant == DG_KERNEL_MFMA) {
        // the MFMA kernel with R |J| M_ref applied to its output tile
        RhsArgs r = rhs_args(p, (const double* const*)workspace, q_bar, false);
        r.post_tiles = p->post_tiles;
        r.jac = p->jac;
        for (int c = 0; c < p->nfields; ++c) r.post_sign[c] = m.sign[c];
        hipLaunchKernelGGL(p->post_kernel, dim3(p->post_grid), dim3(p->kernel.threads), 0,
                           (hipStream_t)stream,
                           r);
        DG_HIP_CHECK(hipGetLastError());
        return DG_OK;
      }
      // the plan's LDS-staged kernel, then R |J| M_ref in place
      if ((rc = dg_rhs(p, DG_PREC_FP64, (const double* const*)workspace, q_bar, stream)) != DG_OK)
        return rc;
      for (int c = 0; c < p->nfields; ++c) {
        m.in[c] = q_bar[c];
        m.out[c] = q_bar[c];
      }
      m.matrix = p->mass;
      m.jac_power = 1;
      hipLaunchKernelGGL(mk, dim3(p->mass_grid), dim3(kBlock), 0, (hipStream_t)stream, m);
      DG_HIP_CHECK(hipGetLastError());
      return DG_OK;
    }
  }

  // the adjoint kernel addresses the fields with 32-bit buffer offsets
  if (fbytes >= ((size_t)1 << 32))
    return fail(DG_ERR_UNSUPPORTED, std::string(name) +
                                        ": each field must stay below 4 GiB for the adjoint kernel");
  if constexpr (!f32) {
    if (p->vjp_mfma_kernel) {
      // Euler on tetrahedra: the mass pass, then the MFMA adjoint kernel
      hipLaunchKernelGGL(mk, dim3(p->mass_grid), dim3(kBlock), 0, (hipStream_t)stream, m);
      DG_HIP_CHECK(hipGetLastError());
      VjpMfmaArgs v{};
      for (int c = 0; c < p->nfields; ++c) {
        v.q[c] = q[c];
        v.mu[c] = workspace[c];
        v.qbar[c] = q_bar[c];
      }
      v.sched = {p->sched_tile, p->part_tile};
      v.vgeo = p->vgeo;
      v.fgeo = p->fgeo;
      v.ext_v = p->vjp_ext;
      v.fmask = p->fmask;
      v.ext_fscale = p->ext_fscale;
      v.face_geo_rx = p->ref_normals_ok ? 1 : 0;
      v.vol_tiles = p->vjp_vol_tiles;
      v.face_tiles = p->vjp_face_tiles;
      v.own_tiles = p->vjp_own_tiles;
      v.nelem = p->nelem;
      v.p0 = p->params[0];
      hipLaunchKernelGGL(p->vjp_mfma_kernel, dim3(p->vjp_mfma_grid), dim3(256), 0,
                         (hipStream_t)stream, v);
      DG_HIP_CHECK(hipGetLastError());
      return DG_OK;
    }
  }
  VjpArgs<T> a{};
  for (int c = 0; c < p->nfields; ++c) {
    a.q[c] = q ? q[c] : nullptr;
    a.mu[c] = workspace[c];
    a.qbar[c] = q_bar[c];
  }
  VjpKernel<T> vk;
  if constexpr (f32)
    vk = p->vjp_kernel32;
  else
    vk = p->vjp_kernel;
  hipLaunchKernelGGL(mk, dim3(p->mass_grid), dim3(kBlock), 0, (hipStream_t)stream, m);
  DG_HIP_CHECK(hipGetLastError());
  a.sched = {p->sched_batch, p->part_batch};
  a.geo = p->geo;
  a.emap = p->emap;
  a.emap_b = p->emap_b;
  a.fmask = p->fmask;
  a.amat_t = p->amat_t;
  a.face_mass = p->face_mass;
  a.ext_fscale = p->ext_fscale;
  a.node_faces = p->node_faces;
  a.nelem = p->nelem;
  a.p0 = p->params[0];
  hipLaunchKernelGGL(vk, dim3(f32 ? p->vjp_grid32 : p->vjp_grid), dim3(kBlock), 0,
                     (hipStream_t)stream, a);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}

// dg_adjoint_space in either storage precision (FP64 arithmetic).
template <typename T>
int adjoint_space_impl(const dg_plan* p, int32_t to_adjoint, const T* const* in, T* const* out,
                       void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (!p->has_vjp)
    return fail(DG_ERR_UNSUPPORTED, "dg_adjoint_space: plan created without mass_inv / face_mass");
  if (p->equation == DG_EQ_EULER)
    return fail(DG_ERR_UNSUPPORTED, "dg_adjoint_space: the Euler operator is not linear");
  int rc;
  if ((rc = check_ptrs((const void* const*)in, p->nfields, "fields_in")) ||
      (rc = check_ptrs((const void* const*)out, p->nfields, "fields_out")))
    return rc;
  MassArgs<T> m{};
  for (int c = 0; c < p->nfields; ++c) {
    m.in[c] = in[c];
    m.out[c] = out[c];
  }
  adjoint_reflection(p, m.sign);
  m.matrix = to_adjoint ? p->mass_inv : p->mass;
  m.jac = p->jac;
  m.jac_power = to_adjoint ? -1 : 1;
  m.nelem = p->nelem;
  MassKernel<T> mk;
  if constexpr (sizeof(T) == 4)
    mk = p->mass_kernel32;
  else
    mk = p->mass_kernel;
  hipLaunchKernelGGL(mk, dim3(p->mass_grid), dim3(kBlock), 0, (hipStream_t)stream, m);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}

template <typename T>
int lserk_adjoint_impl(const dg_plan* p, double a, double b, double dt, T* const* qbar,
                       T* const* sv, const T* const* w, void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  int rc;
  if ((rc = check_ptrs((const void* const*)qbar, p->nfields, "q_bar")) ||
      (rc = check_ptrs((const void* const*)sv, p->nfields, "s")) ||
      (w && (rc = check_ptrs((const void* const*)w, p->nfields, "w"))))
    return rc;
  LserkAdjArgs<T> s{};
  for (int c = 0; c < p->nfields; ++c) {
    s.qbar[c] = qbar[c];
    s.s[c] = sv[c];
    s.w[c] = w ? w[c] : nullptr;
  }
  s.n = p->nelem * p->np;
  s.nf = p->nfields;
  s.a = a;
  s.b = b;
  s.dt = dt;
  hipLaunchKernelGGL(lserk_adjoint_kernel<T>, dim3(p->lserk_grid), dim3(kBlock), 0,
                     (hipStream_t)stream, s);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
}

}  // namespace

extern "C" {

int dg_rhs_vjp(const dg_plan* p, const double* const* q, const double* const* rhs_bar,
               double* const* q_bar, double* const* workspace, void* stream) {
  return rhs_vjp_impl<double>(p, q, rhs_bar, q_bar, workspace, stream, "dg_rhs_vjp");
}

int dg_rhs_vjp_fp32(const dg_plan* p, const float* const* q, const float* const* rhs_bar,
                    float* const* q_bar, float* const* workspace, void* stream) {
  return rhs_vjp_impl<float>(p, q, rhs_bar, q_bar, workspace, stream, "dg_rhs_vjp_fp32");
}

int dg_rhs_vjp_noalloc(const dg_plan* p, const double* const* q, const double* const* rhs_bar,
                       double* const* q_bar, void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (!p->has_vjp || !p->scratch)
    return fail(DG_ERR_UNSUPPORTED,
                "dg_rhs_vjp_noalloc: plan created without mass_inv / face_mass: no adjoint");
  // one field of the workspace per K*Np doubles, each at a 256-byte boundary
  const size_t stride = ((size_t)p->nelem * (size_t)p->np * sizeof(double) + 255) / 256 * 256;
  const size_t need = stride * (size_t)p->nfields;
  if (need > kScratchBytes)
    return fail(DG_ERR_INVALID_ARGUMENT,
                "dg_rhs_vjp_noalloc: the workspace needs " + std::to_string(need) +
                    " bytes (nfields x K x Np doubles), more than the " +
                    std::to_string(kScratchBytes) + "-byte scratch; use dg_rhs_vjp");
  double* ws[kMaxFields];
  for (int c = 0; c < p->nfields; ++c)
    ws[c] = (double*)((unsigned char*)p->scratch + (size_t)c * stride);
  // the mass pass writes every workspace entry before anything reads it, so
  // nothing in the scratch needs clearing
  return rhs_vjp_impl<double>(p, q, rhs_bar, q_bar, ws, stream, "dg_rhs_vjp_noalloc");
}

int dg_adjoint_space(const dg_plan* p, int32_t to_adjoint, const double* const* in,
                     double* const* out, void* stream) {
  return adjoint_space_impl<double>(p, to_adjoint, in, out, stream);
}

int dg_adjoint_space_fp32(const dg_plan* p, int32_t to_adjoint, const float* const* in,
                          float* const* out, void* stream) {
  return adjoint_space_impl<float>(p, to_adjoint, in, out, stream);
}

int dg_lserk_stage_adjoint(const dg_plan* p, double a, double b, double dt, double* const* qbar,
                           double* const* sv, const double* const* w, void* stream) {
  return lserk_adjoint_impl<double>(p, a, b, dt, qbar, sv, w, stream);
}

int dg_lserk_stage_adjoint_fp32(const dg_plan* p, double a, double b, double dt,
                                float* const* qbar, float* const* sv, const float* const* w,
                                void* stream) {
  return lserk_adjoint_impl<float>(p, a, b, dt, qbar, sv, w, stream);
}

}  // extern "C"

// vim: foldmethod=marker
