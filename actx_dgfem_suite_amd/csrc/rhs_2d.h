// Fused right-hand sides on triangles (dim = 2), FP64 arithmetic on FP64 or
// FP32 fields: the three operators of the tetrahedral kernels restricted to
// two space dimensions, for the
// reference's 2D cases (generate_regular_rect_mesh over dim=2 in
// equations/wave.py:248-270, equations/euler.py:150-166,
// equations/maxwell.py:43-66, run with --dims 2,3 in .github/workflows/ci.yml).
//
// One thread per (element, node), batches of kBlock / Np elements staged in
// LDS, the same structure as the LDS-staged tetrahedral kernel: triangles
// carry 2-6x fewer nodes per element, so the data per batch is small and the
// kernel is bound by the same HBM stream (fields in, fields out, geometry and
// the exterior map).
#pragma once

#include "dg_common.h"

namespace dg {

constexpr int kFaces2 = 3;
constexpr int kGeo2 = 13;     // rx (4) + 3 faces x (nx, ny, fscale)

constexpr int np2_of(int order) { return (order + 1) * (order + 2) / 2; }
constexpr int nfp2_of(int order) { return order + 1; }
constexpr int nfields2_of(int eq) { return eq == DG_EQ_EULER ? 4 : 3; }

// {{{ numerical fluxes in 2D: out[c] is the integrand of +face_mass

// wave_flux (equations/wave.py:80-114) with v = (v_x, v_y); mirror walls.
struct Wave2DFlux {
  static constexpr int NF = 3;
  __device__ static void boundary(const double* qm, const double* /*n*/, double* qp) {
    qp[0] = -qm[0];
    qp[1] = qm[1];
    qp[2] = qm[2];
  }
  __device__ static void flux(const double* qm, const double* qp, const double* n,
                              double c, double* out) {
    const double uavg = 0.5 * (qm[0] + qp[0]);
    const double udiff = qp[0] - qm[0];
    const double vavg_n = 0.5 * ((qm[1] + qp[1]) * n[0] + (qm[2] + qp[2]) * n[1]);
    const double vjump = (qp[1] - qm[1]) * n[0] + (qp[2] - qm[2]) * n[1];
    out[0] = c * (vavg_n + 0.5 * udiff);
    const double s = c * (uavg + 0.5 * vjump);
    out[1] = s * n[0];
    out[2] = s * n[1];
  }
};

// Euler, LF flux and reflecting walls; fields (mass, energy, momentum[2]).
struct Euler2DFlux {
  static constexpr int NF = 4;
  __device__ static void boundary(const double* qm, const double* n, double* qp) {
    const double mn = qm[2] * n[0] + qm[3] * n[1];
    qp[0] = qm[0];
    qp[1] = qm[1];
    qp[2] = qm[2] - 2.0 * mn * n[0];
    qp[3] = qm[3] - 2.0 * mn * n[1];
  }
  __device__ static void normal_flux(const double* q, const double* n, double gamma,
                                     double* fn, double* lam) {
    const double rho = q[0], e = q[1];
    const double inv_rho = rcp_nr(rho);
    const double ux = q[2] * inv_rho, uy = q[3] * inv_rho;
    const double p = (gamma - 1.0) * (e - 0.5 * (q[2] * ux + q[3] * uy));
    const double un = ux * n[0] + uy * n[1];
    fn[0] = q[2] * n[0] + q[3] * n[1];
    fn[1] = (e + p) * un;
    fn[2] = q[2] * un + p * n[0];
    fn[3] = q[3] * un + p * n[1];
    *lam = sqrt_nr(ux * ux + uy * uy) + sqrt_nr(gamma * p * inv_rho);
  }
  __device__ static void flux(const double* qm, const double* qp, const double* n,
                              double gamma, double* out) {
    double fm[4], fp[4], lm, lp;
    normal_flux(qm, n, gamma, fm, &lm);
    normal_flux(qp, n, gamma, fp, &lp);
    const double lam = fmax(lm, lp);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      out[c] = -(0.5 * (fm[c] + fp[c]) - 0.5 * lam * (qp[c] - qm[c]));
  }
};

// Maxwell in TM form, fields (E_z, H_x, H_y): the 3D flux with n_z = 0 and
// E_x = E_y = H_z = 0.  With w = n_x [H_y] - n_y [H_x] (the z-component of
// n x [H]) and n x (n x [E]) = -[E_z] z-hat, n x (n x [H]) = (n_y w, -n_x w):
//   flux_Ez = -w/2 - alpha [E_z]/2
//   flux_Hx =  n_y [E_z]/2 + alpha n_y w/2
//   flux_Hy = -n_x [E_z]/2 - alpha n_x w/2
// PEC walls: E_z+ = -E_z-, H+ = H-.  out = -flux.
struct Maxwell2DFlux {
  static constexpr int NF = 3;
  __device__ static void boundary(const double* qm, const double* /*n*/, double* qp) {
    qp[0] = -qm[0];
    qp[1] = qm[1];
    qp[2] = qm[2];
  }
  __device__ static void flux(const double* qm, const double* qp, const double* n,
                              double alpha, double* out) {
    const double dez = qp[0] - qm[0];
    const double w = n[0] * (qp[2] - qm[2]) - n[1] * (qp[1] - qm[1]);
    out[0] = 0.5 * w + 0.5 * alpha * dez;
    out[1] = -0.5 * n[1] * (dez + alpha * w);
    out[2] = 0.5 * n[0] * (dez + alpha * w);
  }
};

template <int EQ> struct Flux2DOf;
template <> struct Flux2DOf<DG_EQ_WAVE> { using type = Wave2DFlux; };
template <> struct Flux2DOf<DG_EQ_EULER> { using type = Euler2DFlux; };
template <> struct Flux2DOf<DG_EQ_MAXWELL> { using type = Maxwell2DFlux; };

// }}}

// Work values per node staged before the contraction: wave w_r = sum_x
// rx[x][r] v_x (2), Euler w_{c,r} = sum_x rx[x][r] F_{c,x} (8); Maxwell
// contracts its fields directly.
constexpr int nwork2_of(int eq) { return eq == DG_EQ_WAVE ? 2 : (eq == DG_EQ_EULER ? 8 : 1); }

// Arguments of the triangle kernel.  T is the storage type of the fields
// (double: the FP64 form; float: the mixed form's FP32 state, with every
// operation still in FP64).  geo rows hold kGeo2 doubles, emap rows 3*Nfp
// entries; dmat_t is (2, Np(j), Np(i)) and lift_t (3*Nfp, Np(i)).
template <typename T>
struct Rhs2DArgs {
  Schedule sched;
  const T* q[kMaxFields];
  T* rhs[kMaxFields];
  const double* geo;
  const int32_t* emap;
  const double* dmat_t;
  const double* lift_t;
  const int32_t* fmask;
  int64_t nelem;
  double p0;
};

template <int EQ, int ORDER, typename T>
__global__ __launch_bounds__(kBlock) void rhs_2d_kernel(Rhs2DArgs<T> a) {
  using Flux = typename Flux2DOf<EQ>::type;
  constexpr int NP = np2_of(ORDER);
  constexpr int NFP = nfp2_of(ORDER);
  constexpr int NF = nfields2_of(EQ);
  constexpr int NFN = kFaces2 * NFP;
  constexpr int EB = kBlock / NP;   // elements per batch
  constexpr int NW = nwork2_of(EQ);

  __shared__ double s_dt[2 * NP * NP];
  __shared__ double s_lt[NFN * NP];
  __shared__ int s_fmask[NFN];
  __shared__ double s_q[NF][EB * NP];
  __shared__ double s_w[NW][EB * NP];
  __shared__ double s_fl[NF][EB * NFN];
  __shared__ double s_geo[EB * kGeo2];

  const int tid = threadIdx.x;
  for (int k = tid; k < 2 * NP * NP; k += kBlock) s_dt[k] = a.dmat_t[k];
  for (int k = tid; k < NFN * NP; k += kBlock) s_lt[k] = a.lift_t[k];
  for (int k = tid; k < NFN; k += kBlock) s_fmask[k] = a.fmask[k];

  const int32_t p_end = a.sched.part[block_xcd() + 1];
  const int32_t p_step = blocks_per_xcd();
  const int el = tid / NP;
  const int i = tid - el * NP;
  const double c0 = a.p0;
  constexpr int R = (EB * NFN + kBlock - 1) / kBlock;   // face-node slots per thread
  // Each thread keeps its rows of D_r^T (node i) in registers, so the volume
  // contraction reads only the (broadcast) staged values from LDS.  Measured
  // on MI355X: -2% to -15% (wave P3 -13%, Maxwell P3 -15%, wave P4 -9%),
  // Euler P3-P4 even; LIFT rows in registers as well were up to 27% slower
  // (registers cost blocks per CU).
  double dr0[NP], dr1[NP];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    dr0[j] = s_dt[j * NP + i];
    dr1[j] = s_dt[NP * NP + j * NP + i];
  }

  for (int32_t pos = a.sched.part[block_xcd()] + block_in_xcd(); pos < p_end; pos += p_step) {
    const int64_t b = sched_at(a.sched.order, pos);
    const int64_t e0 = b * EB;
    const int nel = (int)(a.nelem - e0 < EB ? a.nelem - e0 : EB);
    // exterior indices and values of this batch's face nodes, issued before
    // the staging loads so both are in flight together
    int32_t ext[R];
    double qx[R][NF];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = tid + r * kBlock;
      ext[r] = k < nel * NFN ? a.emap[e0 * NFN + k] : -1;
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < NF; ++c) qx[r][c] = ext[r] >= 0 ? (double)a.q[c][ext[r]] : 0.0;
    __syncthreads();
    {
      // Every staging load in flight before the first LDS store (a
      // load-then-store loop per field waits out one round trip per field);
      // EB * NP <= kBlock, so a thread stages at most one node per field.
      constexpr int RG = (EB * kGeo2 + kBlock - 1) / kBlock;
      const bool own = tid < nel * NP;
      double v[NF], vg[RG];
      if (own) {
#pragma unroll
        for (int c = 0; c < NF; ++c) v[c] = (double)a.q[c][e0 * NP + tid];
      }
#pragma unroll
      for (int r = 0; r < RG; ++r) {
        const int k = tid + r * kBlock;
        if (k < nel * kGeo2) vg[r] = a.geo[e0 * kGeo2 + k];
      }
      if (own) {
#pragma unroll
        for (int c = 0; c < NF; ++c) s_q[c][tid] = v[c];
      }
#pragma unroll
      for (int r = 0; r < RG; ++r) {
        const int k = tid + r * kBlock;
        if (k < nel * kGeo2) s_geo[k] = vg[r];
      }
    }
    __syncthreads();

    const bool active = el < nel;
    if (active) {
      const double* g = &s_geo[el * kGeo2];
      const int node = el * NP + i;
      if constexpr (EQ == DG_EQ_WAVE) {
        const double vx = s_q[1][node], vy = s_q[2][node];
#pragma unroll
        for (int r = 0; r < 2; ++r) s_w[r][node] = g[r] * vx + g[2 + r] * vy;
      } else if constexpr (EQ == DG_EQ_EULER) {
        const double gamma = c0;
        const double rho = s_q[0][node], e = s_q[1][node];
        const double inv_rho = rcp_nr(rho);
        const double m[2] = {s_q[2][node], s_q[3][node]};
        const double u[2] = {m[0] * inv_rho, m[1] * inv_rho};
        const double p = (gamma - 1.0) * (e - 0.5 * (m[0] * u[0] + m[1] * u[1]));
        const double h = (e + p) * inv_rho;
        double f[4][2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          f[0][x] = m[x];
          f[1][x] = m[x] * h;
#pragma unroll
          for (int y = 0; y < 2; ++y) f[2 + y][x] = m[y] * u[x] + (x == y ? p : 0.0);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int r = 0; r < 2; ++r) s_w[c * 2 + r][node] = g[r] * f[c][0] + g[2 + r] * f[c][1];
      }
    }
    // face fluxes scaled by Fscale
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = tid + r * kBlock;
      if (k >= nel * NFN) continue;
      const int fe = k / NFN;
      const int fn = k - fe * NFN;
      const int face = fn / NFP;
      const double* g = &s_geo[fe * kGeo2 + 4 + 3 * face];
      const double n[2] = {g[0], g[1]};
      double qm[NF], qp[NF], out[NF];
      const int vn = fe * NP + s_fmask[fn];
#pragma unroll
      for (int c = 0; c < NF; ++c) qm[c] = s_q[c][vn];
      if (ext[r] >= 0) {
#pragma unroll
        for (int c = 0; c < NF; ++c) qp[c] = qx[r][c];
      } else {
        Flux::boundary(qm, n, qp);
      }
      Flux::flux(qm, qp, n, c0, out);
#pragma unroll
      for (int c = 0; c < NF; ++c) s_fl[c][fe * NFN + fn] = g[2] * out[c];
    }
    __syncthreads();

    if (active) {
      const double* g = &s_geo[el * kGeo2];
      const int base = el * NP;
      double acc[NF];
      if constexpr (EQ == DG_EQ_WAVE) {
        double au = 0.0, t0 = 0.0, t1 = 0.0;
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          const double d0 = dr0[j];
          const double d1 = dr1[j];
          const double uj = s_q[0][base + j];
          au += d0 * s_w[0][base + j] + d1 * s_w[1][base + j];
          t0 += d0 * uj;
          t1 += d1 * uj;
        }
        acc[0] = -c0 * au;
#pragma unroll
        for (int x = 0; x < 2; ++x) acc[1 + x] = -c0 * (g[2 * x] * t0 + g[2 * x + 1] * t1);
      } else if constexpr (EQ == DG_EQ_EULER) {
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = 0.0;
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          const double d0 = dr0[j];
          const double d1 = dr1[j];
#pragma unroll
          for (int c = 0; c < 4; ++c)
            acc[c] += d0 * s_w[2 * c][base + j] + d1 * s_w[2 * c + 1][base + j];
        }
      } else {
        double t[3][2];
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c][0] = t[c][1] = 0.0;
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          const double d0 = dr0[j];
          const double d1 = dr1[j];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double fj = s_q[c][base + j];
            t[c][0] += d0 * fj;
            t[c][1] += d1 * fj;
          }
        }
        // d[c][x] = sum_r rx[x][r] t[c][r]
        double d[3][2];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int x = 0; x < 2; ++x) d[c][x] = g[2 * x] * t[c][0] + g[2 * x + 1] * t[c][1];
        acc[0] = d[2][0] - d[1][1];    // E_z' = dHy/dx - dHx/dy
        acc[1] = -d[0][1];             // H_x' = -dEz/dy
        acc[2] = d[0][0];              // H_y' = dEz/dx
      }
      const int fbase = el * NFN;
#pragma unroll
      for (int fn = 0; fn < NFN; ++fn) {
        const double l = s_lt[fn * NP + i];
#pragma unroll
        for (int c = 0; c < NF; ++c) acc[c] += l * s_fl[c][fbase + fn];
      }
      const int64_t out = e0 * NP + tid;
#pragma unroll
      for (int c = 0; c < NF; ++c) __builtin_nontemporal_store((T)acc[c], a.rhs[c] + out);
    }
  }
}

}  // namespace dg
