// libdgfem.so: DG-FEM right-hand sides (wave, Euler, Maxwell; 3D affine
// tetrahedra, P1-P4) for gfx950, behind the C ABI declared in include/dgfem.h.
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
// with the constant operators stored transposed (transpose_consts_in_einsums.py)
// so that consecutive lanes (consecutive output nodes i) read consecutive LDS
// words.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "dgfem.h"

namespace {

constexpr int kFaces = 4;
constexpr int kGeo = 25;      // rx (9) + 4 faces x (nx, ny, nz, fscale)
constexpr int kBlock = 256;
constexpr int kMaxFields = 6;

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

constexpr int np_of(int order) { return (order + 1) * (order + 2) * (order + 3) / 6; }
constexpr int nfp_of(int order) { return (order + 1) * (order + 2) / 2; }
constexpr int nfields_of(int eq) { return eq == DG_EQ_WAVE ? 4 : (eq == DG_EQ_EULER ? 5 : 6); }

struct InPtrs { const double* p[kMaxFields]; };
struct OutPtrs { double* p[kMaxFields]; };

struct RhsArgs {
  InPtrs q;
  OutPtrs rhs;
  const double* geo;      // (K, 25)
  const int32_t* emap;    // (K, 4*Nfp)
  const double* dmat_t;   // (3, Np(j), Np(i)): transposed weak or strong D
  const double* lift_t;   // (4*Nfp, Np(i))
  const int32_t* fmask;   // (4*Nfp)
  int64_t nelem;
  double p0;
};

// {{{ numerical fluxes: out[c] is the integrand of +face_mass for field c

// wave_flux, equations/wave.py:80-114: c({v}.n + [u]/2, ({u} + [v].n/2) n),
// with [a] = exterior - interior; boundary state u+ = -u-, v+ = v- (:152-156).
struct WaveFlux {
  static constexpr int NF = 4;
  __device__ static void boundary(const double* qm, const double* /*n*/, double* qp) {
    qp[0] = -qm[0];
    qp[1] = qm[1];
    qp[2] = qm[2];
    qp[3] = qm[3];
  }
  __device__ static void flux(const double* qm, const double* qp, const double* n,
                              double c, double* out) {
    const double uavg = 0.5 * (qm[0] + qp[0]);
    const double udiff = qp[0] - qm[0];
    const double vavg_n = 0.5 * ((qm[1] + qp[1]) * n[0] + (qm[2] + qp[2]) * n[1] +
                                 (qm[3] + qp[3]) * n[2]);
    const double vjump = (qp[1] - qm[1]) * n[0] + (qp[2] - qm[2]) * n[1] +
                         (qp[3] - qm[3]) * n[2];
    out[0] = c * (vavg_n + 0.5 * udiff);
    const double s = c * (uavg + 0.5 * vjump);
    out[1] = s * n[0];
    out[2] = s * n[1];
    out[3] = s * n[2];
  }
};

// Euler, local Lax-Friedrichs: F* = {F}.n - lam/2 (q+ - q-), lam = max(|u|+c)
// over both sides; wall state reflects momentum.  The operator subtracts
// face_mass(F*), so out = -F*.  Field order: mass, energy, momentum[3].
struct EulerFlux {
  static constexpr int NF = 5;
  __device__ static void boundary(const double* qm, const double* n, double* qp) {
    const double mn = qm[2] * n[0] + qm[3] * n[1] + qm[4] * n[2];
    qp[0] = qm[0];
    qp[1] = qm[1];
    qp[2] = qm[2] - 2.0 * mn * n[0];
    qp[3] = qm[3] - 2.0 * mn * n[1];
    qp[4] = qm[4] - 2.0 * mn * n[2];
  }
  // normal flux F(q).n and wavespeed |u| + sqrt(gamma p / rho)
  __device__ static void normal_flux(const double* q, const double* n, double gamma,
                                     double* fn, double* lam) {
    const double rho = q[0], e = q[1];
    const double ux = q[2] / rho, uy = q[3] / rho, uz = q[4] / rho;
    const double p = (gamma - 1.0) * (e - 0.5 * (q[2] * ux + q[3] * uy + q[4] * uz));
    const double un = ux * n[0] + uy * n[1] + uz * n[2];
    fn[0] = q[2] * n[0] + q[3] * n[1] + q[4] * n[2];
    fn[1] = (e + p) * un;
    fn[2] = q[2] * un + p * n[0];
    fn[3] = q[3] * un + p * n[1];
    fn[4] = q[4] * un + p * n[2];
    *lam = sqrt(ux * ux + uy * uy + uz * uz) + sqrt(gamma * (p / rho));
  }
  __device__ static void flux(const double* qm, const double* qp, const double* n,
                              double gamma, double* out) {
    double fm[5], fp[5], lm, lp;
    normal_flux(qm, n, gamma, fm, &lm);
    normal_flux(qp, n, gamma, fp, &lp);
    const double lam = fmax(lm, lp);
#pragma unroll
    for (int c = 0; c < 5; ++c)
      out[c] = -(0.5 * (fm[c] + fp[c]) - 0.5 * lam * (qp[c] - qm[c]));
  }
};

// Maxwell (eps = mu = 1), flux_type alpha (Hesthaven & Warburton p. 433):
//   flux_E = -1/2 n x [H] + alpha/2 n x (n x [E])
//   flux_H =  1/2 n x [E] + alpha/2 n x (n x [H])
// PEC walls: E+ = -E-, H+ = H-.  The operator subtracts M^{-1} face_mass(flux).
struct MaxwellFlux {
  static constexpr int NF = 6;
  __device__ static void boundary(const double* qm, const double* /*n*/, double* qp) {
    qp[0] = -qm[0];
    qp[1] = -qm[1];
    qp[2] = -qm[2];
    qp[3] = qm[3];
    qp[4] = qm[4];
    qp[5] = qm[5];
  }
  __device__ static void cross(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  }
  __device__ static void flux(const double* qm, const double* qp, const double* n,
                              double alpha, double* out) {
    double de[3], dh[3], nde[3], ndh[3], nnde[3], nndh[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      de[k] = qp[k] - qm[k];
      dh[k] = qp[3 + k] - qm[3 + k];
    }
    cross(n, de, nde);
    cross(n, dh, ndh);
    cross(n, nde, nnde);
    cross(n, ndh, nndh);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      out[k] = -(-0.5 * ndh[k] + 0.5 * alpha * nnde[k]);
      out[3 + k] = -(0.5 * nde[k] + 0.5 * alpha * nndh[k]);
    }
  }
};

template <int EQ> struct FluxOf;
template <> struct FluxOf<DG_EQ_WAVE> { using type = WaveFlux; };
template <> struct FluxOf<DG_EQ_EULER> { using type = EulerFlux; };
template <> struct FluxOf<DG_EQ_MAXWELL> { using type = MaxwellFlux; };

// }}}

// {{{ fused FP64 RHS kernel

// Work values per node staged in LDS before the contraction: wave keeps
// w_r = sum_x rx[x][r] v_x (3), Euler keeps w_{c,r} = sum_x rx[x][r] F_{c,x}
// (15); Maxwell contracts the fields themselves.
constexpr int nwork_of(int eq) { return eq == DG_EQ_WAVE ? 3 : (eq == DG_EQ_EULER ? 15 : 1); }

template <int EQ, int ORDER>
__global__ __launch_bounds__(kBlock) void rhs_fp64_kernel(RhsArgs a) {
  using Flux = typename FluxOf<EQ>::type;
  constexpr int NP = np_of(ORDER);
  constexpr int NFP = nfp_of(ORDER);
  constexpr int NF = nfields_of(EQ);
  constexpr int NFN = kFaces * NFP;
  constexpr int EB = kBlock / NP;   // elements per batch
  constexpr int NW = nwork_of(EQ);

  __shared__ double s_dt[3 * NP * NP];
  __shared__ double s_lt[NFN * NP];
  __shared__ int s_fmask[NFN];
  __shared__ double s_q[NF][EB * NP];
  __shared__ double s_w[NW][EB * NP];
  __shared__ double s_fl[NF][EB * NFN];
  __shared__ double s_geo[EB * kGeo];

  const int tid = threadIdx.x;
  for (int k = tid; k < 3 * NP * NP; k += kBlock) s_dt[k] = a.dmat_t[k];
  for (int k = tid; k < NFN * NP; k += kBlock) s_lt[k] = a.lift_t[k];
  for (int k = tid; k < NFN; k += kBlock) s_fmask[k] = a.fmask[k];

  // XCD-aware batch ranges: blocks are dealt round-robin to the 8 XCDs, so give
  // each XCD one contiguous run of batches (neighbour gathers then stay in that
  // XCD's L2).
  const int64_t nbatch = (a.nelem + EB - 1) / EB;
  const int nblk = gridDim.x;
  const int xcd_blocks = (nblk + 7) / 8;
  const int lblk = (blockIdx.x % 8) * xcd_blocks + blockIdx.x / 8;
  const int64_t per = (nbatch + (int64_t)xcd_blocks * 8 - 1) / ((int64_t)xcd_blocks * 8);
  const int64_t b_begin = (int64_t)lblk * per;
  const int64_t b_end = b_begin + per < nbatch ? b_begin + per : nbatch;

  const int el = tid / NP;
  const int i = tid - el * NP;
  const double c0 = a.p0;

  for (int64_t b = b_begin; b < b_end; ++b) {
    const int64_t e0 = b * EB;
    const int nel = (int)(a.nelem - e0 < EB ? a.nelem - e0 : EB);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      const double* src = a.q.p[c] + e0 * NP;
      for (int k = tid; k < nel * NP; k += kBlock) s_q[c][k] = src[k];
    }
    for (int k = tid; k < nel * kGeo; k += kBlock) s_geo[k] = a.geo[e0 * kGeo + k];
    __syncthreads();

    const bool active = el < nel;
    // pointwise work values
    if (active) {
      const double* g = &s_geo[el * kGeo];
      const int node = el * NP + i;
      if constexpr (EQ == DG_EQ_WAVE) {
        const double vx = s_q[1][node], vy = s_q[2][node], vz = s_q[3][node];
#pragma unroll
        for (int r = 0; r < 3; ++r) s_w[r][node] = g[r] * vx + g[3 + r] * vy + g[6 + r] * vz;
      } else if constexpr (EQ == DG_EQ_EULER) {
        const double gamma = c0;
        const double rho = s_q[0][node], e = s_q[1][node];
        const double m[3] = {s_q[2][node], s_q[3][node], s_q[4][node]};
        const double u[3] = {m[0] / rho, m[1] / rho, m[2] / rho};
        const double p = (gamma - 1.0) * (e - 0.5 * (m[0] * u[0] + m[1] * u[1] + m[2] * u[2]));
        // flux tensor F[c][x]
        double f[5][3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          f[0][x] = m[x];
          f[1][x] = m[x] * (e + p) / rho;
#pragma unroll
          for (int y = 0; y < 3; ++y) f[2 + y][x] = m[y] * m[x] / rho + (x == y ? p : 0.0);
        }
#pragma unroll
        for (int c = 0; c < 5; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r)
            s_w[c * 3 + r][node] = g[r] * f[c][0] + g[3 + r] * f[c][1] + g[6 + r] * f[c][2];
      }
    }
    // face fluxes, scaled by Fscale
    for (int k = tid; k < nel * NFN; k += kBlock) {
      const int fe = k / NFN;
      const int fn = k - fe * NFN;
      const int face = fn / NFP;
      const double* g = &s_geo[fe * kGeo + 9 + 4 * face];
      const double n[3] = {g[0], g[1], g[2]};
      double qm[NF], qp[NF], out[NF];
      const int vn = fe * NP + s_fmask[fn];
#pragma unroll
      for (int c = 0; c < NF; ++c) qm[c] = s_q[c][vn];
      const int32_t ext = a.emap[(e0 + fe) * NFN + fn];
      if (ext >= 0) {
#pragma unroll
        for (int c = 0; c < NF; ++c) qp[c] = a.q.p[c][ext];
      } else {
        Flux::boundary(qm, n, qp);
      }
      Flux::flux(qm, qp, n, c0, out);
#pragma unroll
      for (int c = 0; c < NF; ++c) s_fl[c][fe * NFN + fn] = g[3] * out[c];
    }
    __syncthreads();

    if (active) {
      const double* g = &s_geo[el * kGeo];
      const int base = el * NP;
      double acc[NF];
      if constexpr (EQ == DG_EQ_WAVE) {
        double au = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll 4
        for (int j = 0; j < NP; ++j) {
          const double d0 = s_dt[j * NP + i];
          const double d1 = s_dt[NP * NP + j * NP + i];
          const double d2 = s_dt[2 * NP * NP + j * NP + i];
          const double uj = s_q[0][base + j];
          au += d0 * s_w[0][base + j] + d1 * s_w[1][base + j] + d2 * s_w[2][base + j];
          t0 += d0 * uj;
          t1 += d1 * uj;
          t2 += d2 * uj;
        }
        acc[0] = -c0 * au;
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[1 + x] = -c0 * (g[3 * x] * t0 + g[3 * x + 1] * t1 + g[3 * x + 2] * t2);
      } else if constexpr (EQ == DG_EQ_EULER) {
#pragma unroll
        for (int c = 0; c < 5; ++c) acc[c] = 0.0;
#pragma unroll 2
        for (int j = 0; j < NP; ++j) {
          const double d0 = s_dt[j * NP + i];
          const double d1 = s_dt[NP * NP + j * NP + i];
          const double d2 = s_dt[2 * NP * NP + j * NP + i];
#pragma unroll
          for (int c = 0; c < 5; ++c)
            acc[c] += d0 * s_w[3 * c][base + j] + d1 * s_w[3 * c + 1][base + j] +
                      d2 * s_w[3 * c + 2][base + j];
        }
      } else {
        double t[6][3];
#pragma unroll
        for (int c = 0; c < 6; ++c) t[c][0] = t[c][1] = t[c][2] = 0.0;
#pragma unroll 2
        for (int j = 0; j < NP; ++j) {
          const double d0 = s_dt[j * NP + i];
          const double d1 = s_dt[NP * NP + j * NP + i];
          const double d2 = s_dt[2 * NP * NP + j * NP + i];
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            const double fj = s_q[c][base + j];
            t[c][0] += d0 * fj;
            t[c][1] += d1 * fj;
            t[c][2] += d2 * fj;
          }
        }
        // physical derivative d_x f = sum_r rx[x][r] t_r
        double d[6][3];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
          for (int x = 0; x < 3; ++x)
            d[c][x] = g[3 * x] * t[c][0] + g[3 * x + 1] * t[c][1] + g[3 * x + 2] * t[c][2];
        // E_t = curl H, H_t = -curl E
        acc[0] = d[5][1] - d[4][2];
        acc[1] = d[3][2] - d[5][0];
        acc[2] = d[4][0] - d[3][1];
        acc[3] = -(d[2][1] - d[1][2]);
        acc[4] = -(d[0][2] - d[2][0]);
        acc[5] = -(d[1][0] - d[0][1]);
      }
      // LIFT
      const int fbase = el * NFN;
#pragma unroll 4
      for (int fn = 0; fn < NFN; ++fn) {
        const double l = s_lt[fn * NP + i];
#pragma unroll
        for (int c = 0; c < NF; ++c) acc[c] += l * s_fl[c][fbase + fn];
      }
      const int64_t out = e0 * NP + tid;
#pragma unroll
      for (int c = 0; c < NF; ++c) a.rhs.p[c][out] = acc[c];
    }
  }
}

// }}}

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
      const double r = s.a * res[k] + s.dt * rhs[k];
      res[k] = r;
      q[k] += s.b * r;
    }
  }
}

// }}}

using RhsKernel = void (*)(RhsArgs);

template <int EQ>
RhsKernel pick_fp64(int order) {
  switch (order) {
    case 1: return rhs_fp64_kernel<EQ, 1>;
    case 2: return rhs_fp64_kernel<EQ, 2>;
    case 3: return rhs_fp64_kernel<EQ, 3>;
    case 4: return rhs_fp64_kernel<EQ, 4>;
    default: return nullptr;
  }
}

RhsKernel pick_kernel(int eq, int order) {
  switch (eq) {
    case DG_EQ_WAVE: return pick_fp64<DG_EQ_WAVE>(order);
    case DG_EQ_EULER: return pick_fp64<DG_EQ_EULER>(order);
    case DG_EQ_MAXWELL: return pick_fp64<DG_EQ_MAXWELL>(order);
    default: return nullptr;
  }
}

}  // namespace

struct dg_plan {
  int equation = 0;
  int order = 0;
  int np = 0;
  int nfp = 0;
  int nfields = 0;
  int64_t nelem = 0;
  double params[4] = {0, 0, 0, 0};
  int device = 0;
  int grid = 0;          // persistent grid of the RHS kernel
  int lserk_grid = 0;
  RhsKernel kernel = nullptr;
  double* geo = nullptr;
  int32_t* emap = nullptr;
  double* dmat_t = nullptr;
  double* lift_t = nullptr;
  int32_t* fmask = nullptr;
};

namespace {

void free_plan(dg_plan* p) {
  if (!p) return;
  hipFree(p->geo);
  hipFree(p->emap);
  hipFree(p->dmat_t);
  hipFree(p->lift_t);
  hipFree(p->fmask);
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

int check_ptrs(const void* const* ptrs, int n, const char* what) {
  if (!ptrs) return fail(DG_ERR_INVALID_ARGUMENT, std::string(what) + " is NULL");
  for (int k = 0; k < n; ++k)
    if (!ptrs[k]) return fail(DG_ERR_INVALID_ARGUMENT, std::string(what) + "[" + std::to_string(k) + "] is NULL");
  return DG_OK;
}

}  // namespace

extern "C" {

const char* dg_last_error_string(void) { return g_last_error.c_str(); }

int dg_plan_create(const dg_plan_desc* d, dg_plan** out) {
  g_last_error.clear();
  if (!d || !out) return fail(DG_ERR_INVALID_ARGUMENT, "NULL descriptor or output");
  *out = nullptr;
  if (d->equation < DG_EQ_WAVE || d->equation > DG_EQ_MAXWELL)
    return fail(DG_ERR_UNSUPPORTED, "unknown equation id " + std::to_string(d->equation));
  if (d->order < 1 || d->order > 4)
    return fail(DG_ERR_UNSUPPORTED, "order " + std::to_string(d->order) + " not supported (1-4)");
  if (d->nelements < 1) return fail(DG_ERR_INVALID_ARGUMENT, "nelements must be >= 1");
  const int np = np_of(d->order), nfp = nfp_of(d->order);
  if ((int64_t)d->nelements * np >= (int64_t)INT32_MAX ||
      (int64_t)d->nelements * kFaces * nfp >= (int64_t)INT32_MAX)
    return fail(DG_ERR_UNSUPPORTED, "mesh too large for 32-bit node indices");
  if (!d->d_weak || !d->d_strong || !d->lift || !d->face_nodes || !d->geometry || !d->ext_map)
    return fail(DG_ERR_INVALID_ARGUMENT, "NULL operator or mesh array in descriptor");

  const int64_t k = d->nelements;
  const int nfn = kFaces * nfp;
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

  dg_plan* p = new (std::nothrow) dg_plan;
  if (!p) return fail(DG_ERR_OUT_OF_MEMORY, "host allocation failed");
  p->equation = d->equation;
  p->order = d->order;
  p->np = np;
  p->nfp = nfp;
  p->nfields = nfields_of(d->equation);
  p->nelem = k;
  for (int t = 0; t < 4; ++t) p->params[t] = d->params[t];
  p->kernel = pick_kernel(d->equation, d->order);

  int rc;
  if (hipGetDevice(&p->device) != hipSuccess) {
    free_plan(p);
    return fail(DG_ERR_HIP, "hipGetDevice failed");
  }
  // transposed operators: dmat_t[r][j][i] = D[r][i][j]; lift_t[fn][i] = LIFT[i][fn]
  const double* dsrc = d->equation == DG_EQ_MAXWELL ? d->d_strong : d->d_weak;
  std::vector<double> dt(3 * np * np), lt((size_t)nfn * np);
  for (int r = 0; r < 3; ++r)
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) dt[(r * np + j) * np + i] = dsrc[(r * np + i) * np + j];
  for (int i = 0; i < np; ++i)
    for (int fn = 0; fn < nfn; ++fn) lt[fn * np + i] = d->lift[i * nfn + fn];
  if ((rc = upload(&p->dmat_t, dt.data(), dt.size())) ||
      (rc = upload(&p->lift_t, lt.data(), lt.size())) ||
      (rc = upload(&p->fmask, d->face_nodes, (size_t)nfn)) ||
      (rc = upload(&p->geo, d->geometry, (size_t)k * kGeo)) ||
      (rc = upload(&p->emap, d->ext_map, (size_t)k * nfn))) {
    free_plan(p);
    return rc;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) {
    free_plan(p);
    return fail(DG_ERR_HIP, "hipGetDeviceProperties failed");
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)p->kernel, kBlock, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int eb = kBlock / np;
  const int64_t nbatch = (k + eb - 1) / eb;
  int64_t grid = (int64_t)prop.multiProcessorCount * per_cu;
  if (grid > nbatch) grid = nbatch;
  // the kernel's XCD remap needs a multiple of 8 blocks; spare blocks idle
  grid = (grid + 7) / 8 * 8;
  p->grid = (int)grid;
  p->lserk_grid = prop.multiProcessorCount * 8;
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

int dg_rhs(const dg_plan* p, int32_t precision, const double* const* in, double* const* out,
           void* stream) {
  if (!p) return fail(DG_ERR_INVALID_ARGUMENT, "NULL plan");
  if (precision != DG_PREC_FP64)
    return fail(DG_ERR_UNSUPPORTED, "precision " + std::to_string(precision) + " not supported");
  int rc;
  if ((rc = check_ptrs((const void* const*)in, p->nfields, "fields_in")) ||
      (rc = check_ptrs((const void* const*)out, p->nfields, "fields_out")))
    return rc;
  RhsArgs a{};
  for (int c = 0; c < p->nfields; ++c) {
    a.q.p[c] = in[c];
    a.rhs.p[c] = out[c];
  }
  a.geo = p->geo;
  a.emap = p->emap;
  a.dmat_t = p->dmat_t;
  a.lift_t = p->lift_t;
  a.fmask = p->fmask;
  a.nelem = p->nelem;
  a.p0 = p->params[0];
  hipLaunchKernelGGL(p->kernel, dim3(p->grid), dim3(kBlock), 0, (hipStream_t)stream, a);
  DG_HIP_CHECK(hipGetLastError());
  return DG_OK;
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

// vim: foldmethod=marker
