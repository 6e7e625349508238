// Adjoint (vector-Jacobian product) of the DG right-hand side:
//   q_bar = (dL/dq)^T lam   for rhs = L(q)
// Exact transpose of the forward kernels' arithmetic for wave and Maxwell
// (linear), and of the Euler operator linearised at q (flux Jacobians
// recomputed per node).  This file holds the per-element mass operator
// (mass_apply_kernel) and the adjoint kernel (rhs_vjp_kernel).  The FP64
// adjoint of wave and Maxwell on tetrahedra does not use the adjoint kernel:
// the host runs it as R M L R M^{-1} lam on the forward kernel, with the
// M multiplication fused into the MFMA kernel (dgfem.hip, rhs_vjp_impl;
// rhs_mfma.h, POST).
//
// The adjoint kernel path has two passes.  The first forms mu = M^{-1} lam per element.  Everything the
// second pass needs is a function of mu:
//   volume:  D_r^T lam = (D_r^T M) mu = A_r mu       (A_r = S_r in weak form)
//   surface: the forward term rhs_c[e,i] += sum_fn LIFT[i][fn] fs_e(f) Phi_c
//            with LIFT = M^{-1} E transposes to the face cotangent
//            psi_e(f, a) = fs_e(f) sum_b Mf_f[b][a] mu_e[fmask_f[b]]
//            (E couples only the nodes of one face).
// Surface transpose without atomics: the exterior map is an involution, so the
// contribution of the neighbour's flux to this element's face node is a
// GATHER of the neighbour's psi at the matching face node, which in this
// element's face numbering is
//   psi_e'(f, a) = fs_e'(f') sum_b Mf_f[b][a] mu_e'(ext(e, f, b))
// (the face mass of one physical face is the same in both numberings), and
//   q_bar(fmask[fn]) += JM^T psi_e + [boundary: B^T JP^T psi_e]
//                     + JP(q_e'(ext), q_e(fn), -n)^T psi_e'.
// For the conservative wave and Euler fluxes Phi(a, b, -n) = -Phi(b, a, n),
// so the last term is -JM^T psi_e' and one flux VJP with psi_e - psi_e'
// covers both.  Each element's face nodes are then summed into its volume
// nodes (vertex nodes lie on three faces, edge nodes on two).
//
// Euler's LF wavespeed max(lam_M, lam_P) takes the subgradient that splits
// exact ties 0.5 / 0.5 (the convention of JAX's maximum); ties occur at wall
// faces, where both branches have the same derivative.
//
// State, cotangents and results are float64 or float32 (the FP32 adjoint);
// the arithmetic is FP64 in both.
#pragma once

#include "dg_common.h"
#include "rhs_2d.h"

namespace dg {

template <typename T>
struct VjpArgs {
  Schedule sched;                // batches of kBlock / Np elements
  const T* q[kMaxFields];        // state (Euler only; NULL for wave / Maxwell)
  const T* mu[kMaxFields];       // M^{-1} lam per element (first pass)
  T* qbar[kMaxFields];           // result
  const double* geo;             // (K, 25)
  const int32_t* emap;           // (K, 4*Nfp)
  // the exterior map with neighbours inside the element's batch pointing into
  // the batch's staged state (bit 30 set, batch-local element in bits 8-15,
  // node in bits 0-7)
  const int32_t* emap_b;
  const int32_t* fmask;          // (4*Nfp)
  const double* amat_t;          // (3, Np(j), Np(i)): A_r[i][j] = (D_r^T M)[i][j]
  const double* face_mass;       // (4, Nfp, Nfp) Mf[f][b][a]
  const double* ext_fscale;      // (K, 4) Fscale of the neighbour across each face
  const int32_t* node_faces;     // (Np, 3) face-node indices of each volume node, -1 padded
  int64_t nelem;
  double p0;
};

// Per-element mass operator on every field: out_c = sign_c * J^p * A in_c,
// with A the reference M^{-1} or M (symmetric), J the element Jacobian (p = -1
// or +1; no scaling without a Jacobian table).  out may alias in.
template <typename T>
struct MassArgs {
  const T* in[kMaxFields];
  T* out[kMaxFields];
  const double* matrix;          // (Np, Np), symmetric
  const double* jac;             // (K) |J| per element, or NULL
  int32_t jac_power;             // -1: divide by J, +1: multiply by J
  double sign[kMaxFields];
  int64_t nelem;
};

// {{{ flux VJPs: given psi (cotangent of the flux 'out' of FluxOf<EQ>::flux),
// gM = (d out / d qM)^T psi, gP = (d out / d qP)^T psi; bcT: g <- B^T g for the
// boundary state qP = B qM.

struct WaveAdj {
  static constexpr int NF = 4;
  static constexpr bool CONSERVATIVE = true;
  __device__ static void vjp_m(const double*, const double*, const double* n, double c,
                               const double* psi, double* gM) {
    const double pvn = psi[1] * n[0] + psi[2] * n[1] + psi[3] * n[2];
    gM[0] = 0.5 * c * (-psi[0] + pvn);
#pragma unroll
    for (int k = 0; k < 3; ++k) gM[1 + k] = 0.5 * c * (psi[0] - pvn) * n[k];
  }
  __device__ static void vjp(const double*, const double*, const double* n, double c,
                             const double* psi, double* gM, double* gP) {
    const double pvn = psi[1] * n[0] + psi[2] * n[1] + psi[3] * n[2];
    gM[0] = 0.5 * c * (-psi[0] + pvn);
    gP[0] = 0.5 * c * (psi[0] + pvn);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gM[1 + k] = 0.5 * c * (psi[0] - pvn) * n[k];
      gP[1 + k] = 0.5 * c * (psi[0] + pvn) * n[k];
    }
  }
  __device__ static void bcT(const double*, double* g) { g[0] = -g[0]; }
};

struct MaxwellAdj {
  static constexpr int NF = 6;
  static constexpr bool CONSERVATIVE = false;   // strong form: flux minus own trace
  __device__ static void vjp_m(const double*, const double*, const double*, double,
                               const double*, double*) {}
  __device__ static void cross(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  }
  __device__ static void vjp(const double*, const double*, const double* n, double alpha,
                             const double* psi, double* gM, double* gP) {
    double ne[3], nh[3], nne[3], nnh[3];
    cross(n, psi, ne);        // n x psi_E
    cross(n, psi + 3, nh);    // n x psi_H
    cross(n, ne, nne);
    cross(n, nh, nnh);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double de = -0.5 * alpha * nne[k] + 0.5 * nh[k];
      const double dh = -0.5 * ne[k] - 0.5 * alpha * nnh[k];
      gP[k] = de;
      gP[3 + k] = dh;
      gM[k] = -de;
      gM[3 + k] = -dh;
    }
  }
  __device__ static void bcT(const double*, double* g) {
    g[0] = -g[0];
    g[1] = -g[1];
    g[2] = -g[2];
  }
};

struct EulerAdj {
  static constexpr int NF = 5;
  static constexpr bool CONSERVATIVE = true;
  // g += (d (F(q).n) / d q)^T psi, n need not be a unit vector
  __device__ static void nflux_vjp(const double* q, const double* n, double gamma,
                                   const double* psi, double* g) {
    const double rho = q[0], e = q[1];
    const double m[3] = {q[2], q[3], q[4]};
    const double inv = rcp_nr(rho);
    const double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    const double p = (gamma - 1.0) * (e - 0.5 * m2 * inv);
    const double un = (m[0] * n[0] + m[1] * n[1] + m[2] * n[2]) * inv;
    const double pmn = psi[2] * n[0] + psi[3] * n[1] + psi[4] * n[2];
    const double pmm = psi[2] * m[0] + psi[3] * m[1] + psi[4] * m[2];
    const double Q = psi[1] * (e + p) + pmm;
    const double R = psi[1] * un + pmn;
    g[0] += R * (gamma - 1.0) * 0.5 * m2 * inv * inv - Q * un * inv;
    g[1] += psi[1] * un + (gamma - 1.0) * R;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      g[2 + k] += psi[0] * n[k] + Q * inv * n[k] + un * psi[2 + k] - (gamma - 1.0) * R * m[k] * inv;
  }
  // wavespeed |u| + c and its gradient
  __device__ static double speed(const double* q, double gamma, double* ds) {
    const double rho = q[0], e = q[1];
    const double inv = rcp_nr(rho);
    const double u[3] = {q[2] * inv, q[3] * inv, q[4] * inv};
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double p = (gamma - 1.0) * (e - 0.5 * rho * uu);
    const double c = sqrt_nr(gamma * p * inv);
    const double au = sqrt_nr(uu);
    const double dp_drho = (gamma - 1.0) * 0.5 * uu;
    // two reciprocals instead of a division per term
    const double k = 0.5 * gamma * inv * rcp_nr(c);
    const double iau = au > 0.0 ? inv * rcp_nr(au) : 0.0;
    ds[0] = (au > 0.0 ? -au * inv : 0.0) + k * (dp_drho - p * inv);
    ds[1] = k * (gamma - 1.0);
#pragma unroll
    for (int j = 0; j < 3; ++j) ds[2 + j] = u[j] * iau - k * (gamma - 1.0) * u[j];
    return au + c;
  }
  // gM (and gP when NEED_P) of the LF flux
  template <bool NEED_P>
  __device__ static void vjp_impl(const double* qM, const double* qP, const double* n,
                                  double gamma, const double* psi, double* gM, double* gP) {
    double dsM[5], dsP[5];
    const double lM = speed(qM, gamma, dsM);
    const double lP = speed(qP, gamma, dsP);
    const double lam = fmax(lM, lP);
    const double wM = lM > lP ? 1.0 : (lM < lP ? 0.0 : 0.5);
    const double wP = 1.0 - wM;
    double jump = 0.0;
#pragma unroll
    for (int c = 0; c < 5; ++c) jump += psi[c] * (qP[c] - qM[c]);
    double hpsi[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      hpsi[c] = -0.5 * psi[c];
      gM[c] = -0.5 * lam * psi[c] + 0.5 * jump * wM * dsM[c];
      if constexpr (NEED_P) gP[c] = 0.5 * lam * psi[c] + 0.5 * jump * wP * dsP[c];
    }
    nflux_vjp(qM, n, gamma, hpsi, gM);
    if constexpr (NEED_P) nflux_vjp(qP, n, gamma, hpsi, gP);
  }
  __device__ static void vjp(const double* qM, const double* qP, const double* n, double gamma,
                             const double* psi, double* gM, double* gP) {
    vjp_impl<true>(qM, qP, n, gamma, psi, gM, gP);
  }
  __device__ static void vjp_m(const double* qM, const double* qP, const double* n, double gamma,
                               const double* psi, double* gM) {
    vjp_impl<false>(qM, qP, n, gamma, psi, gM, nullptr);
  }
  __device__ static void bcT(const double* n, double* g) {
    const double gn = g[2] * n[0] + g[3] * n[1] + g[4] * n[2];
    g[2] -= 2.0 * gn * n[0];
    g[3] -= 2.0 * gn * n[1];
    g[4] -= 2.0 * gn * n[2];
  }
};

template <int EQ> struct AdjOf;
template <> struct AdjOf<DG_EQ_WAVE> { using type = WaveAdj; };
template <> struct AdjOf<DG_EQ_EULER> { using type = EulerAdj; };
template <> struct AdjOf<DG_EQ_MAXWELL> { using type = MaxwellAdj; };

// {{{ the same for the triangle fluxes of rhs_2d.h

// Wave2DFlux: out = c({v}.n + [u]/2, ({u} + [v].n/2) n) with two components.
struct Wave2DAdj {
  static constexpr int NF = 3;
  static constexpr bool CONSERVATIVE = true;
  __device__ static void vjp_m(const double*, const double*, const double* n, double c,
                               const double* psi, double* gM) {
    const double pvn = psi[1] * n[0] + psi[2] * n[1];
    gM[0] = 0.5 * c * (-psi[0] + pvn);
#pragma unroll
    for (int k = 0; k < 2; ++k) gM[1 + k] = 0.5 * c * (psi[0] - pvn) * n[k];
  }
  __device__ static void vjp(const double*, const double*, const double* n, double c,
                             const double* psi, double* gM, double* gP) {
    const double pvn = psi[1] * n[0] + psi[2] * n[1];
    gM[0] = 0.5 * c * (-psi[0] + pvn);
    gP[0] = 0.5 * c * (psi[0] + pvn);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      gM[1 + k] = 0.5 * c * (psi[0] - pvn) * n[k];
      gP[1 + k] = 0.5 * c * (psi[0] + pvn) * n[k];
    }
  }
  __device__ static void bcT(const double*, double* g) { g[0] = -g[0]; }
};

// Maxwell2DFlux (TM): out depends on the jump qP - qM only, through
// d = [E_z] and w = n_x [H_y] - n_y [H_x]:
//   out = (w/2 + alpha d/2, -n_y (d + alpha w)/2, n_x (d + alpha w)/2).
// With a = n_x psi_Hy - n_y psi_Hx:
//   gP = (alpha psi_Ez/2 + a/2, -n_y (psi_Ez + alpha a)/2, n_x (psi_Ez + alpha a)/2), gM = -gP.
struct Maxwell2DAdj {
  static constexpr int NF = 3;
  static constexpr bool CONSERVATIVE = false;
  __device__ static void vjp_m(const double*, const double*, const double*, double,
                               const double*, double*) {}
  __device__ static void vjp(const double*, const double*, const double* n, double alpha,
                             const double* psi, double* gM, double* gP) {
    const double av = n[0] * psi[2] - n[1] * psi[1];
    const double t = psi[0] + alpha * av;
    gP[0] = 0.5 * alpha * psi[0] + 0.5 * av;
    gP[1] = -0.5 * n[1] * t;
    gP[2] = 0.5 * n[0] * t;
#pragma unroll
    for (int c = 0; c < 3; ++c) gM[c] = -gP[c];
  }
  __device__ static void bcT(const double*, double* g) { g[0] = -g[0]; }
};

// Euler2DFlux: LF flux with fields (mass, energy, momentum[2]).
struct Euler2DAdj {
  static constexpr int NF = 4;
  static constexpr bool CONSERVATIVE = true;
  // g += (d (F(q).n) / d q)^T psi, n need not be a unit vector
  __device__ static void nflux_vjp(const double* q, const double* n, double gamma,
                                   const double* psi, double* g) {
    const double rho = q[0], e = q[1];
    const double m[2] = {q[2], q[3]};
    const double inv = rcp_nr(rho);
    const double m2 = m[0] * m[0] + m[1] * m[1];
    const double p = (gamma - 1.0) * (e - 0.5 * m2 * inv);
    const double un = (m[0] * n[0] + m[1] * n[1]) * inv;
    const double pmn = psi[2] * n[0] + psi[3] * n[1];
    const double pmm = psi[2] * m[0] + psi[3] * m[1];
    const double Q = psi[1] * (e + p) + pmm;
    const double R = psi[1] * un + pmn;
    g[0] += R * (gamma - 1.0) * 0.5 * m2 * inv * inv - Q * un * inv;
    g[1] += psi[1] * un + (gamma - 1.0) * R;
#pragma unroll
    for (int k = 0; k < 2; ++k)
      g[2 + k] += psi[0] * n[k] + Q * inv * n[k] + un * psi[2 + k] - (gamma - 1.0) * R * m[k] * inv;
  }
  // wavespeed |u| + c and its gradient
  __device__ static double speed(const double* q, double gamma, double* ds) {
    const double rho = q[0], e = q[1];
    const double inv = rcp_nr(rho);
    const double u[2] = {q[2] * inv, q[3] * inv};
    const double uu = u[0] * u[0] + u[1] * u[1];
    const double p = (gamma - 1.0) * (e - 0.5 * rho * uu);
    const double c = sqrt_nr(gamma * p * inv);
    const double au = sqrt_nr(uu);
    const double dp_drho = (gamma - 1.0) * 0.5 * uu;
    const double k = 0.5 * gamma * inv * rcp_nr(c);
    const double iau = au > 0.0 ? inv * rcp_nr(au) : 0.0;
    ds[0] = (au > 0.0 ? -au * inv : 0.0) + k * (dp_drho - p * inv);
    ds[1] = k * (gamma - 1.0);
#pragma unroll
    for (int j = 0; j < 2; ++j) ds[2 + j] = u[j] * iau - k * (gamma - 1.0) * u[j];
    return au + c;
  }
  template <bool NEED_P>
  __device__ static void vjp_impl(const double* qM, const double* qP, const double* n,
                                  double gamma, const double* psi, double* gM, double* gP) {
    double dsM[4], dsP[4];
    const double lM = speed(qM, gamma, dsM);
    const double lP = speed(qP, gamma, dsP);
    const double lam = fmax(lM, lP);
    const double wM = lM > lP ? 1.0 : (lM < lP ? 0.0 : 0.5);
    const double wP = 1.0 - wM;
    double jump = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) jump += psi[c] * (qP[c] - qM[c]);
    double hpsi[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      hpsi[c] = -0.5 * psi[c];
      gM[c] = -0.5 * lam * psi[c] + 0.5 * jump * wM * dsM[c];
      if constexpr (NEED_P) gP[c] = 0.5 * lam * psi[c] + 0.5 * jump * wP * dsP[c];
    }
    nflux_vjp(qM, n, gamma, hpsi, gM);
    if constexpr (NEED_P) nflux_vjp(qP, n, gamma, hpsi, gP);
  }
  __device__ static void vjp(const double* qM, const double* qP, const double* n, double gamma,
                             const double* psi, double* gM, double* gP) {
    vjp_impl<true>(qM, qP, n, gamma, psi, gM, gP);
  }
  __device__ static void vjp_m(const double* qM, const double* qP, const double* n, double gamma,
                               const double* psi, double* gM) {
    vjp_impl<false>(qM, qP, n, gamma, psi, gM, nullptr);
  }
  __device__ static void bcT(const double* n, double* g) {
    const double gn = g[2] * n[0] + g[3] * n[1];
    g[2] -= 2.0 * gn * n[0];
    g[3] -= 2.0 * gn * n[1];
  }
};

template <int EQ> struct Adj2DOf;
template <> struct Adj2DOf<DG_EQ_WAVE> { using type = Wave2DAdj; };
template <> struct Adj2DOf<DG_EQ_EULER> { using type = Euler2DAdj; };
template <> struct Adj2DOf<DG_EQ_MAXWELL> { using type = Maxwell2DAdj; };

// }}}

// }}}

// {{{ per-element mass operator: batches of kBlock / Np elements; every
// field of a batch is staged in LDS with coalesced loads, then one thread per
// (element, node) forms its row of A in and stores it coalesced.  A
// persistent grid of several blocks per CU keeps loads of many batches in
// flight.  In place is safe: a block reads its whole batch into LDS (two
// barriers) before any thread of it writes, and batches are disjoint.

template <int NP, int NF, typename T>
__global__ __launch_bounds__(kBlock) void mass_apply_kernel(MassArgs<T> a) {
  constexpr int EB = kBlock / NP;
  __shared__ double s_m[NP * NP];     // s_m[j * NP + i] = A[i][j] (= A[j][i])
  __shared__ double s_l[NF][EB * NP];
  const int tid = threadIdx.x;
  for (int k = tid; k < NP * NP; k += kBlock) s_m[k] = a.matrix[k];
  const int el = tid / NP, i = tid - el * NP;
  const int64_t nbatch = (a.nelem + EB - 1) / EB;
  // Each thread keeps its row of A in registers, so the contraction reads
  // only the (broadcast) staged values from LDS: with A from LDS too the
  // kernel is bound by LDS reads, two per multiply-add (measured on MI355X:
  // wave P3 58 -> 35 us per pass).
  double arow[NP];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NP; ++j) arow[j] = s_m[j * NP + i];
  for (int64_t b = blockIdx.x; b < nbatch; b += gridDim.x) {
    const int64_t e0 = b * EB;
    const int nel = (int)(a.nelem - e0 < EB ? a.nelem - e0 : EB);
    const bool own = tid < nel * NP;   // EB * NP <= kBlock: one node per thread
    // all fields' loads in flight before the first LDS store
    double v[NF];
    if (own) {
#pragma unroll
      for (int c = 0; c < NF; ++c) v[c] = (double)a.in[c][e0 * NP + tid];
    }
    double scale = 1.0;
    if (a.jac && el < nel) scale = a.jac_power > 0 ? a.jac[e0 + el] : 1.0 / a.jac[e0 + el];
    __syncthreads();
    if (own) {
#pragma unroll
      for (int c = 0; c < NF; ++c) s_l[c][tid] = v[c];
    }
    __syncthreads();
    if (el < nel) {
#pragma unroll
      for (int c = 0; c < NF; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < NP; ++j)
          acc += arow[j] * s_l[c][el * NP + j];
        a.out[c][e0 * NP + tid] = (T)(a.sign[c] * (scale * acc));
      }
    }
  }
}

// }}}

// {{{ adjoint kernel: one thread per (element, node), element batches in LDS

// Shapes of the adjoint kernel on tetrahedra (DIM 3) and triangles (DIM 2):
// geometry rows hold rx (DIM * DIM), then per face DIM normal components and
// Fscale.
template <int DIM>
struct AdjShape {
  static constexpr int FACES = DIM + 1;
  static constexpr int GEO = DIM == 3 ? kGeo : kGeo2;
  static constexpr int RX = DIM * DIM;
  static constexpr int FACE_STRIDE = DIM + 1;
  static constexpr int np(int order) { return DIM == 3 ? np_of(order) : np2_of(order); }
  static constexpr int nfp(int order) { return DIM == 3 ? nfp_of(order) : nfp2_of(order); }
  static constexpr int nfields(int eq) { return DIM == 3 ? nfields_of(eq) : nfields2_of(eq); }
};

template <int EQ, int DIM> struct AdjFluxOf;
template <int EQ> struct AdjFluxOf<EQ, 3> {
  using adj = typename AdjOf<EQ>::type;
  using flux = typename FluxOf<EQ>::type;
};
template <int EQ> struct AdjFluxOf<EQ, 2> {
  using adj = typename Adj2DOf<EQ>::type;
  using flux = typename Flux2DOf<EQ>::type;
};

template <int EQ, int ORDER, typename T, int DIM = 3>
__global__ __launch_bounds__(kBlock) void rhs_vjp_kernel(VjpArgs<T> a) {
  using Sh = AdjShape<DIM>;
  using Adj = typename AdjFluxOf<EQ, DIM>::adj;
  using Flux = typename AdjFluxOf<EQ, DIM>::flux;
  constexpr int NP = Sh::np(ORDER);
  constexpr int NFP = Sh::nfp(ORDER);
  constexpr int NF = Sh::nfields(EQ);
  constexpr int NFACES = Sh::FACES;
  constexpr int GEO = Sh::GEO;
  constexpr int NFN = NFACES * NFP;
  constexpr int EB = kBlock / NP;
  constexpr bool NEEDS_Q = EQ == DG_EQ_EULER;
  constexpr int NQ = NEEDS_Q ? NF : 1;
  constexpr int NW = EQ == DG_EQ_WAVE ? DIM : 1;

  __shared__ double s_a[DIM * NP * NP];      // A_r[i][j] at [r][j][i]
  __shared__ double s_mf[NFACES * NFP * NFP];
  __shared__ int s_fmask[NFN];
  __shared__ int s_nf[NP * 3];
  __shared__ double s_mu[NF][EB * NP];
  __shared__ double s_q[NQ][EB * NP];
  __shared__ double s_w[NW][EB * NP];
  __shared__ double s_mux[NF][EB * NFN];     // mu at exterior nodes
  // q at exterior nodes (Euler), overwritten in place (same thread, same
  // slot) by the face-node contributions to q_bar
  __shared__ double s_x[NF][EB * NFN];
  __shared__ double s_geo[EB * GEO];
  __shared__ double s_xfs[EB * NFACES];
  // the batch's exterior map, kept from the gathers where it costs no block
  // per CU (else the face phase reads it again from global memory)
  constexpr int LDS_BASE =
      8 * (DIM * NP * NP + NFACES * NFP * NFP +
           (NF + (NEEDS_Q ? NQ : 0) + (EQ == DG_EQ_WAVE ? NW : 0)) * EB * NP + 2 * NF * EB * NFN +
           EB * GEO + EB * NFACES) +
      4 * (NFN + 3 * NP);
  constexpr bool MAP_LDS = kLdsPerCu / (LDS_BASE + 4 * EB * NFN) == kLdsPerCu / LDS_BASE;
  __shared__ int32_t s_ext[MAP_LDS ? EB * NFN : 1];

  const int tid = threadIdx.x;
  for (int k = tid; k < DIM * NP * NP; k += kBlock) s_a[k] = a.amat_t[k];
  for (int k = tid; k < NFACES * NFP * NFP; k += kBlock) s_mf[k] = a.face_mass[k];
  for (int k = tid; k < NFN; k += kBlock) s_fmask[k] = a.fmask[k];
  for (int k = tid; k < NP * 3; k += kBlock) s_nf[k] = a.node_faces[k];

  const int32_t p_end = a.sched.part[block_xcd() + 1];
  const int32_t p_step = blocks_per_xcd();
  const int el = tid / NP;
  const int i = tid - el * NP;
  const double c0 = a.p0;

  // The exterior gathers of the next batch are issued into registers while
  // this batch is processed, so their latency overlaps the face and volume
  // phases.  R face-node slots per thread cover the batch's EB * NFN nodes.
  constexpr int R = (EB * NFN + kBlock - 1) / kBlock;
  constexpr int RG = (EB * GEO + kBlock - 1) / kBlock;   // geometry values per thread
  constexpr int GR = RG < 4 ? RG : 4;
  static_assert(EB * NP <= kBlock && EB * NFACES <= kBlock, "one staged node per thread");
  int32_t pf_ext[R];
  double pf_mu[R][NF];
  double pf_q[R][NQ];
  auto batch_nel = [&](int64_t bb) {
    const int64_t e = bb * EB;
    return (int)(a.nelem - e < EB ? a.nelem - e : EB);
  };
  // Neighbours inside the batch from its staged state (measured on MI355X:
  // FP32 Maxwell P4 -7.5%, wave P3 -2%; Euler, whose kernel is bound by the
  // flux VJPs' arithmetic, +4% to +10% for the extra barrier, so not there)
  constexpr bool IB = EQ != DG_EQ_EULER;
  auto issue_map = [&](int64_t bb) {
    const int n = batch_nel(bb) * NFN;
    const int32_t* map = IB ? a.emap_b : a.emap;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = tid + r * kBlock;
      pf_ext[r] = k < n ? map[bb * EB * NFN + k] : -1;
    }
  };
  // Exterior values of neighbours outside the batch, by buffer loads whose
  // offset is past the field's end for every other lane (boundary nodes and
  // neighbours inside the batch, taken from the staged state below): the
  // buffer unit answers those without a memory access.  The plan keeps each
  // field below 4 GiB.
  const uint32_t field_bytes = (uint32_t)(a.nelem * NP * (int64_t)sizeof(T));
  __amdgpu_buffer_rsrc_t rs_mu[NF], rs_q[NQ];
#pragma unroll
  for (int c = 0; c < NF; ++c)
    rs_mu[c] = __builtin_amdgcn_make_buffer_rsrc((void*)a.mu[c], (short)0, field_bytes, kBufferWord3);
  if constexpr (NEEDS_Q)
#pragma unroll
    for (int c = 0; c < NQ; ++c)
      rs_q[c] = __builtin_amdgcn_make_buffer_rsrc((void*)a.q[c], (short)0, field_bytes, kBufferWord3);
  auto buffer_value = [&](__amdgpu_buffer_rsrc_t rs, uint32_t off) -> double {
    if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
    else
      return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
  };
  auto issue_gathers = [&]() {
    if constexpr (!IB) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < NF; ++c) {
          pf_mu[r][c] = pf_ext[r] >= 0 ? (double)a.mu[c][pf_ext[r]] : 0.0;
          if constexpr (NEEDS_Q) pf_q[r][c] = pf_ext[r] >= 0 ? (double)a.q[c][pf_ext[r]] : 0.0;
        }
      return;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool glob = pf_ext[r] >= 0 && pf_ext[r] < kInTile;
      const uint32_t off = glob ? (uint32_t)pf_ext[r] * (uint32_t)sizeof(T) : field_bytes;
#pragma unroll
      for (int c = 0; c < NF; ++c) {
        pf_mu[r][c] = buffer_value(rs_mu[c], off);
        if constexpr (NEEDS_Q) pf_q[r][c] = buffer_value(rs_q[c], off);
      }
    }
  };

  int32_t pos = a.sched.part[block_xcd()] + block_in_xcd();
  if (pos < p_end) {
    issue_map(sched_at(a.sched.order, pos));
    issue_gathers();
  }
  for (; pos < p_end; pos += p_step) {
    const int64_t b = sched_at(a.sched.order, pos);
    const int64_t e0 = b * EB;
    const int nel = batch_nel(b);
    __syncthreads();
    // this batch's exterior values, gathered during the previous batch
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = tid + r * kBlock;
      if (k < nel * NFN) {
        if constexpr (MAP_LDS) s_ext[k] = pf_ext[r];
#pragma unroll
        for (int c = 0; c < NF; ++c) {
          s_mux[c][k] = pf_mu[r][c];
          if constexpr (NEEDS_Q) s_x[c][k] = pf_q[r][c];
        }
      }
    }
    const bool has_next = pos + p_step < p_end;
    int32_t cur_ext[R];
#pragma unroll
    for (int r = 0; r < R; ++r) cur_ext[r] = pf_ext[r];
    if (has_next) issue_map(sched_at(a.sched.order, pos + p_step));
    // The staging loads of a round are issued together before their LDS
    // stores (a load-then-store loop per field would wait out one memory
    // round trip per field): mu and Fscale, then the geometry, then q
    // (Euler), so that the registers of the rounds do not add up.
    // EB * NP <= kBlock and EB * NFACES <= kBlock, so each thread stages at
    // most one node of each field and one face's Fscale.
    const bool own = tid < nel * NP;
    {
      double v_mu[NF], v_xfs = 0.0;
      if (own) {
#pragma unroll
        for (int c = 0; c < NF; ++c) v_mu[c] = (double)a.mu[c][e0 * NP + tid];
      }
      if (tid < nel * NFACES) v_xfs = a.ext_fscale[e0 * NFACES + tid];
      if (own) {
#pragma unroll
        for (int c = 0; c < NF; ++c) s_mu[c][tid] = v_mu[c];
      }
      if (tid < nel * NFACES) s_xfs[tid] = v_xfs;
    }
    // geometry: at most GR values per thread in flight per round
#pragma unroll 1
    for (int r0 = 0; r0 < RG; r0 += GR) {
      double v_geo[GR];
#pragma unroll
      for (int r = 0; r < GR; ++r) {
        const int k = tid + (r0 + r) * kBlock;
        if (r0 + r < RG && k < nel * GEO) v_geo[r] = a.geo[e0 * GEO + k];
      }
#pragma unroll
      for (int r = 0; r < GR; ++r) {
        const int k = tid + (r0 + r) * kBlock;
        if (r0 + r < RG && k < nel * GEO) s_geo[k] = v_geo[r];
      }
    }
    if constexpr (NEEDS_Q) {
      double v_q[NQ];
      if (own) {
#pragma unroll
        for (int c = 0; c < NQ; ++c) v_q[c] = (double)a.q[c][e0 * NP + tid];
#pragma unroll
        for (int c = 0; c < NQ; ++c) s_q[c][tid] = v_q[c];
      }
    }
    if (has_next) issue_gathers();
    __syncthreads();
    // exterior values of neighbours inside the batch, from its staged state
    if constexpr (IB) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = tid + r * kBlock;
      if (k < nel * NFN && cur_ext[r] >= kInTile) {
        const int lofs = ((cur_ext[r] >> 8) & 255) * NP + (cur_ext[r] & 255);
#pragma unroll
        for (int c = 0; c < NF; ++c) {
          s_mux[c][k] = s_mu[c][lofs];
          if constexpr (NEEDS_Q) s_x[c][k] = s_q[c][lofs];
        }
      }
    }
    __syncthreads();
    }

    const bool active = el < nel;
    if constexpr (EQ == DG_EQ_WAVE) {
      if (active) {
        const double* g = &s_geo[el * GEO];
        const int node = el * NP + i;
        if constexpr (DIM == 3) {
#pragma unroll
          for (int r = 0; r < 3; ++r)
            s_w[r][node] = g[r] * s_mu[1][node] + g[3 + r] * s_mu[2][node] + g[6 + r] * s_mu[3][node];
        } else {
#pragma unroll
          for (int r = 0; r < 2; ++r) s_w[r][node] = g[r] * s_mu[1][node] + g[2 + r] * s_mu[2][node];
        }
      }
    }

    // face-node contributions
    for (int k = tid; k < nel * NFN; k += kBlock) {
      const int fe = k / NFN;
      const int fn = k - fe * NFN;
      const int f = fn / NFP;
      const int aa = fn - f * NFP;
      const double* geo = &s_geo[fe * GEO + Sh::RX + Sh::FACE_STRIDE * f];
      double n[DIM];
#pragma unroll
      for (int x = 0; x < DIM; ++x) n[x] = geo[x];
      const double fs = geo[DIM];
      const double xfs = s_xfs[fe * NFACES + f];
      double psi[NF], psin[NF];
#pragma unroll
      for (int c = 0; c < NF; ++c) {
        double so = 0.0, sn = 0.0;
        for (int bb = 0; bb < NFP; ++bb) {
          const double m = s_mf[(f * NFP + bb) * NFP + aa];
          so += m * s_mu[c][fe * NP + s_fmask[f * NFP + bb]];
          sn += m * s_mux[c][fe * NFN + f * NFP + bb];
        }
        psi[c] = fs * so;
        psin[c] = xfs * sn;
      }
      const int32_t ext = MAP_LDS ? s_ext[k] : a.emap[(e0 + fe) * NFN + fn];
      double qm[NF], qp[NF], gM[NF], gP[NF];
#pragma unroll
      for (int c = 0; c < NF; ++c) {
        qm[c] = NEEDS_Q ? s_q[c < NQ ? c : 0][fe * NP + s_fmask[fn]] : 0.0;
        qp[c] = NEEDS_Q ? s_x[c][k] : 0.0;
      }
      if (ext < 0) {
        if constexpr (NEEDS_Q) Flux::boundary(qm, n, qp);
        Adj::vjp(qm, qp, n, c0, psi, gM, gP);
        Adj::bcT(n, gP);
#pragma unroll
        for (int c = 0; c < NF; ++c) gM[c] += gP[c];
      } else if constexpr (Adj::CONSERVATIVE) {
        double pd[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) pd[c] = psi[c] - psin[c];
        Adj::vjp_m(qm, qp, n, c0, pd, gM);
      } else {
        double nm[DIM];
#pragma unroll
        for (int x = 0; x < DIM; ++x) nm[x] = -n[x];
        Adj::vjp(qm, qp, n, c0, psi, gM, gP);
        double hM[NF], hP[NF];
        // the neighbour's flux: its interior is qp, its exterior is our qm
        Adj::vjp(qp, qm, nm, c0, psin, hM, hP);
#pragma unroll
        for (int c = 0; c < NF; ++c) gM[c] += hP[c];
      }
#pragma unroll
      for (int c = 0; c < NF; ++c) s_x[c][k] = gM[c];
    }
    __syncthreads();

    if (active) {
      const double* g = &s_geo[el * GEO];
      const int base = el * NP;
      double acc[NF];
      if constexpr (DIM == 2) {
        // t_{c,r} = A_r mu_c (wave: the weak divergence's transpose on w)
        if constexpr (EQ == DG_EQ_WAVE) {
          double au = 0.0, t0 = 0.0, t1 = 0.0;
          for (int j = 0; j < NP; ++j) {
            const double d0 = s_a[j * NP + i];
            const double d1 = s_a[NP * NP + j * NP + i];
            const double mu = s_mu[0][base + j];
            au += d0 * s_w[0][base + j] + d1 * s_w[1][base + j];
            t0 += d0 * mu;
            t1 += d1 * mu;
          }
          acc[0] = -c0 * au;
#pragma unroll
          for (int x = 0; x < 2; ++x) acc[1 + x] = -c0 * (g[2 * x] * t0 + g[2 * x + 1] * t1);
        } else {
          double t[NF][2];
#pragma unroll
          for (int c = 0; c < NF; ++c) t[c][0] = t[c][1] = 0.0;
          for (int j = 0; j < NP; ++j) {
            const double d0 = s_a[j * NP + i];
            const double d1 = s_a[NP * NP + j * NP + i];
#pragma unroll
            for (int c = 0; c < NF; ++c) {
              const double l = s_mu[c][base + j];
              t[c][0] += d0 * l;
              t[c][1] += d1 * l;
            }
          }
          // d[c][x] = sum_r rx[x][r] t[c][r]
          double d[NF][2];
#pragma unroll
          for (int c = 0; c < NF; ++c)
#pragma unroll
            for (int x = 0; x < 2; ++x) d[c][x] = g[2 * x] * t[c][0] + g[2 * x + 1] * t[c][1];
          if constexpr (EQ == DG_EQ_EULER) {
            double q[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              q[c] = s_q[c][base + i];
              acc[c] = 0.0;
            }
#pragma unroll
            for (int x = 0; x < 2; ++x) {
              double fb[4];
#pragma unroll
              for (int c = 0; c < 4; ++c) fb[c] = d[c][x];
              const double ex[2] = {x == 0 ? 1.0 : 0.0, x == 1 ? 1.0 : 0.0};
              Euler2DAdj::nflux_vjp(q, ex, c0, fb, acc);
            }
          } else {
            // TM transpose: E_z_bar = dx^T lam_Hy - dy^T lam_Hx,
            // H_x_bar = -dy^T lam_Ez, H_y_bar = dx^T lam_Ez
            acc[0] = d[2][0] - d[1][1];
            acc[1] = -d[0][1];
            acc[2] = d[0][0];
          }
        }
      } else if constexpr (EQ == DG_EQ_WAVE) {
        // u_bar = -c sum_r A_r (sum_x rx[x][r] mu_vx); v_x_bar = -c sum_r rx[x][r] A_r mu_u
        double au = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
        for (int j = 0; j < NP; ++j) {
          const double d0 = s_a[j * NP + i];
          const double d1 = s_a[NP * NP + j * NP + i];
          const double d2 = s_a[2 * NP * NP + j * NP + i];
          const double mu = s_mu[0][base + j];
          au += d0 * s_w[0][base + j] + d1 * s_w[1][base + j] + d2 * s_w[2][base + j];
          t0 += d0 * mu;
          t1 += d1 * mu;
          t2 += d2 * mu;
        }
        acc[0] = -c0 * au;
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[1 + x] = -c0 * (g[3 * x] * t0 + g[3 * x + 1] * t1 + g[3 * x + 2] * t2);
      } else if constexpr (EQ == DG_EQ_EULER) {
        // T_{c,r} = A_r mu_c; Fbar_{c,x} = sum_r rx[x][r] T_{c,r}; q_bar = sum_x (dF_x/dq)^T Fbar_x
        double t[5][3];
#pragma unroll
        for (int c = 0; c < 5; ++c) t[c][0] = t[c][1] = t[c][2] = 0.0;
        for (int j = 0; j < NP; ++j) {
          const double d0 = s_a[j * NP + i];
          const double d1 = s_a[NP * NP + j * NP + i];
          const double d2 = s_a[2 * NP * NP + j * NP + i];
#pragma unroll
          for (int c = 0; c < 5; ++c) {
            const double l = s_mu[c][base + j];
            t[c][0] += d0 * l;
            t[c][1] += d1 * l;
            t[c][2] += d2 * l;
          }
        }
        double q[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          q[c] = s_q[c][base + i];
          acc[c] = 0.0;
        }
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          double fb[5];
#pragma unroll
          for (int c = 0; c < 5; ++c) fb[c] = g[3 * x] * t[c][0] + g[3 * x + 1] * t[c][1] + g[3 * x + 2] * t[c][2];
          const double ex[3] = {x == 0 ? 1.0 : 0.0, x == 1 ? 1.0 : 0.0, x == 2 ? 1.0 : 0.0};
          EulerAdj::nflux_vjp(q, ex, c0, fb, acc);
        }
      } else {
        // H_bar_a = sum eps_{oxa} d_x^T lam_{E_o}; E_bar_a = -sum eps_{oxa} d_x^T lam_{H_o}
        double t[6][3];
#pragma unroll
        for (int c = 0; c < 6; ++c) t[c][0] = t[c][1] = t[c][2] = 0.0;
        for (int j = 0; j < NP; ++j) {
          const double d0 = s_a[j * NP + i];
          const double d1 = s_a[NP * NP + j * NP + i];
          const double d2 = s_a[2 * NP * NP + j * NP + i];
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            const double l = s_mu[c][base + j];
            t[c][0] += d0 * l;
            t[c][1] += d1 * l;
            t[c][2] += d2 * l;
          }
        }
        double d[6][3];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
          for (int x = 0; x < 3; ++x) d[c][x] = g[3 * x] * t[c][0] + g[3 * x + 1] * t[c][1] + g[3 * x + 2] * t[c][2];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] = 0.0;
#pragma unroll
        for (int comp = 0; comp < 3; ++comp)
#pragma unroll
          for (int o = 0; o < 3; ++o)
#pragma unroll
            for (int x = 0; x < 3; ++x) {
              const int s = (o == x || x == comp || o == comp) ? 0 : (((x - o + 3) % 3 == 1) ? 1 : -1);
              if (s == 0) continue;
              acc[3 + comp] += s * d[o][x];       // H_bar from lam_E
              acc[comp] -= s * d[3 + o][x];       // E_bar from lam_H
            }
      }
      // face-node contributions of this node
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int fn = s_nf[i * 3 + s];
        if (fn >= 0) {
#pragma unroll
          for (int c = 0; c < NF; ++c) acc[c] += s_x[c][el * NFN + fn];
        }
      }
      const int64_t out = e0 * NP + tid;
#pragma unroll
      for (int c = 0; c < NF; ++c) __builtin_nontemporal_store((T)acc[c], a.qbar[c] + out);
    }
  }
}

// }}}

}  // namespace dg
