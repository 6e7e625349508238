// Shared definitions of the DG right-hand-side kernels: sizes, argument
// blocks and the numerical-flux functors of the three operators.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dgfem.h"

namespace dg {

constexpr int kFaces = 4;
constexpr int kGeo = 25;      // rx (9) + 4 faces x (nx, ny, nz, fscale)
constexpr int kBlock = 256;
constexpr int kMaxFields = 6;
constexpr int kLdsPerCu = 160 * 1024;   // LDS bytes per CU (gfx950)

constexpr int np_of(int order) { return (order + 1) * (order + 2) * (order + 3) / 6; }
constexpr int nfp_of(int order) { return (order + 1) * (order + 2) / 2; }
constexpr int nfields_of(int eq) { return eq == DG_EQ_WAVE ? 4 : (eq == DG_EQ_EULER ? 5 : 6); }

struct InPtrs { const double* p[kMaxFields]; };
struct OutPtrs { double* p[kMaxFields]; };

// Work schedule of a kernel whose units are runs of consecutive elements
// (tiles or batches): `order` lists the units in processing order and
// `part[x] .. part[x+1]` is the slice of `order` that XCD x works through.
// Workgroups are dispatched round-robin over the 8 XCDs, so XCD x runs blocks
// x, x + 8, ...; all workers of one XCD step through their slice together
// (worker s takes positions s, s + nworkers, ...), which keeps the units in
// flight on an XCD -- and the neighbours their exterior gathers touch --
// inside a compact region of the mesh that fits that XCD's L2.
struct Schedule {
  const int32_t* order;
  const int32_t* part;   // 9 offsets
};

struct RhsArgs {
  Schedule sched;
  InPtrs q;
  OutPtrs rhs;
  const double* geo;      // (K, 25)
  const int32_t* emap;    // (K, 4*Nfp)
  const double* dmat_t;   // (3, Np(j), Np(i)): transposed weak or strong D
  const double* lift_t;   // (4*Nfp, Np(i))
  const int32_t* fmask;   // (4*Nfp)
  const double* op_tiles;    // MFMA A tiles of D (wave: -c D): [r][mblock][kstep][64]
  const double* lift_tiles;  // MFMA A tiles of LIFT: [mblock][kstep][64]
  // MFMA kernel, tile-ordered copies of geometry and exterior map (16-element
  // tiles, the last one padded with its final element):
  const double* vgeo;        // [tile][9][16]: rx of element n of the tile
  const double* fgeo;        // [tile][face*16 + n][4]: n_x, n_y, n_z, Fscale
  // 1 when every face satisfies Fscale n = rx^T c_f (affine images of the
  // reference element), so that the kernel may form its face geometry from rx
  int32_t face_geo_rx;
  const int32_t* ext_t;      // [tile][chunk][item][64]: emap in flux-item order
  const int32_t* ext_d;      // [tile][kstep][64]: emap in the direct flux order
  // MFMA kernel with the mass post-multiplication (the linear operators'
  // adjoint): out_c = post_sign[c] * |J| * M_ref * rhs_c per element
  const double* post_tiles;  // MFMA A tiles of M_ref: [mblock][kstep][64]
  const double* jac;         // (K) |J| per element
  double post_sign[kMaxFields];
  // MFMA kernel with the LSERK4 stage fused into its output (STAGE), from
  // the input state q: res = a * res + dt * rhs (res not read when a == 0);
  // q_out = beta * q + b * res.  q_out must not be q (other tiles gather
  // from q).  Forward stage: beta = 1; reverse stage in the adjoint
  // variables: res is q_hat, a = 1, q is s.
  OutPtrs stage_res;
  OutPtrs stage_qout;
  double stage_a, stage_b, stage_dt, stage_beta;
  int64_t nelem;
  double p0;
};

// Fourth word of a raw buffer resource on gfx950 (32-bit data format).
constexpr int kBufferWord3 = 0x00020000;

// Exterior-map entries with this bit set point into a kernel's LDS copy of the
// state of its current unit (tile or batch) instead of into the state arrays.
constexpr int32_t kInTile = 1 << 30;

// The schedule's unit at a wave-uniform position, loaded through the scalar
// cache.  A vector load would be counted together with the wave's earlier
// output stores, so waiting for it would wait for those stores too.
__device__ inline int32_t sched_at(const int32_t* order, int32_t pos) {
  typedef const __attribute__((address_space(4))) int32_t* const_i32;
  return ((const_i32)order)[__builtin_amdgcn_readfirstlane(pos)];
}

// This block's XCD and its index among the blocks of that XCD.  Needs
// gridDim.x % 8 == 0.
__device__ inline int block_xcd() { return blockIdx.x & 7; }
__device__ inline int block_in_xcd() { return blockIdx.x >> 3; }
__device__ inline int blocks_per_xcd() { return gridDim.x >> 3; }

// {{{ numerical fluxes: out[c] is the integrand of +face_mass for field c

// wave_flux, equations/wave.py:80-114: c({v}.n + [u]/2, ({u} + [v].n/2) n),
// with [a] = exterior - interior; boundary state u+ = -u-, v+ = v- (:152-156).
struct WaveFlux {
  static constexpr int NF = 4;
  // the boundary state is the interior trace with per-field signs (u+ = -u-)
  __host__ __device__ static constexpr int mirror_sign(int c) { return c == 0 ? -1 : 1; }
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

// 1/x and sqrt(x) for the positive, normal, moderate values of the Euler
// state (densities, pressures, squared speeds): the hardware estimates
// v_rcp_f64 and v_rsq_f64 refined by two Newton-Raphson steps (1/x), and by
// one Goldschmidt step and a Newton correction (sqrt), in place of the
// correctly rounded sequences with their range scaling.  Measured on gfx950
// over 4.2M log-uniform samples in [6e-6, 1.6e5]: equal to the IEEE results
// bit for bit, and the Euler RHS bit-identical; Euler P2 RHS -5%, its adjoint
// -2% to -9% across P1-P4.  sqrt_nr(0) = 0, and a negative argument (a state
// with negative pressure) gives NaN, as sqrt does (the LF flux's fmax then
// takes the other side's wavespeed, as it does with sqrt).
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ inline double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  const double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return x == 0.0 ? 0.0 : g;
}

// Euler, local Lax-Friedrichs: F* = {F}.n - lam/2 (q+ - q-), lam = max(|u|+c)
// over both sides; wall state reflects momentum.  The operator subtracts
// face_mass(F*), so out = -F*.  Field order: mass, energy, momentum[3].
struct EulerFlux {
  static constexpr int NF = 5;
  // the wall state depends on the normal: not a per-field sign
  __host__ __device__ static constexpr int mirror_sign(int) { return 0; }
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
    const double inv_rho = rcp_nr(rho);
    const double ux = q[2] * inv_rho, uy = q[3] * inv_rho, uz = q[4] * inv_rho;
    const double p = (gamma - 1.0) * (e - 0.5 * (q[2] * ux + q[3] * uy + q[4] * uz));
    const double un = ux * n[0] + uy * n[1] + uz * n[2];
    fn[0] = q[2] * n[0] + q[3] * n[1] + q[4] * n[2];
    fn[1] = (e + p) * un;
    fn[2] = q[2] * un + p * n[0];
    fn[3] = q[3] * un + p * n[1];
    fn[4] = q[4] * un + p * n[2];
    *lam = sqrt_nr(ux * ux + uy * uy + uz * uz) + sqrt_nr(gamma * p * inv_rho);
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
  // PEC: E+ = -E-, H+ = H-
  __host__ __device__ static constexpr int mirror_sign(int c) { return c < 3 ? -1 : 1; }
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


}  // namespace dg
