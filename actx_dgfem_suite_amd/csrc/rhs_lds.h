// Fused RHS kernel with one thread per (element, node), element batches and
// the transposed operators staged in LDS: the baseline the MFMA kernel is
// measured against (plan variant "lds").
#pragma once

#include "dg_common.h"

namespace dg {

// {{{ fused FP64 RHS kernel, LDS-staged (baseline variant)

// Work values per node staged in LDS before the contraction: wave keeps
// w_r = sum_x rx[x][r] v_x (3), Euler keeps w_{c,r} = sum_x rx[x][r] F_{c,x}
// (15); Maxwell contracts the fields themselves.
constexpr int nwork_of(int eq) { return eq == DG_EQ_WAVE ? 3 : (eq == DG_EQ_EULER ? 15 : 1); }

template <int EQ, int ORDER>
__global__ __launch_bounds__(kBlock) void rhs_lds_kernel(RhsArgs a) {
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

  const int32_t p_end = a.sched.part[block_xcd() + 1];
  const int32_t p_step = blocks_per_xcd();
  const int el = tid / NP;
  const int i = tid - el * NP;
  const double c0 = a.p0;
  // Euler and Maxwell P2 keep each thread's rows of D_r^T in registers, so
  // the contraction reads only the (broadcast) staged values from LDS.
  // Measured on MI355X: Maxwell P2 -5%, Euler P2 -3%; wave P1-P2 +3-13% and
  // P1/P3 even (registers cost blocks per CU).
  constexpr bool ROWS = EQ != DG_EQ_WAVE && ORDER == 2;
  constexpr int UNROLL_J = ROWS ? NP : 2;
  double dr[3][ROWS ? NP : 1];
  if constexpr (ROWS) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int j = 0; j < NP; ++j) dr[r][j] = s_dt[r * NP * NP + j * NP + i];
  }

  // Exterior values of the next batch are gathered into registers during
  // this batch's volume phase; R face-node slots per thread cover a batch.
  constexpr int R = (EB * NFN + kBlock - 1) / kBlock;
  int32_t pf_ext[R];
  double pf_q[R][NF];
  auto batch_nel = [&](int64_t bb) {
    const int64_t e = bb * EB;
    return (int)(a.nelem - e < EB ? a.nelem - e : EB);
  };
  auto issue_map = [&](int64_t bb) {
    const int n = batch_nel(bb) * NFN;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = tid + r * kBlock;
      pf_ext[r] = k < n ? a.emap[bb * EB * NFN + k] : -1;
    }
  };
  auto issue_gathers = [&]() {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < NF; ++c) pf_q[r][c] = pf_ext[r] >= 0 ? a.q.p[c][pf_ext[r]] : 0.0;
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
    const bool has_next = pos + p_step < p_end;
    __syncthreads();
    {
      // Every staging load in flight before the first LDS store (a
      // load-then-store loop per field waits out one round trip per field);
      // EB * NP <= kBlock, so a thread stages at most one node per field.
      constexpr int RG = (EB * kGeo + kBlock - 1) / kBlock;
      const bool own = tid < nel * NP;
      double v[NF], vg[RG];
      if (own) {
#pragma unroll
        for (int c = 0; c < NF; ++c) v[c] = a.q.p[c][e0 * NP + tid];
      }
#pragma unroll
      for (int r = 0; r < RG; ++r) {
        const int k = tid + r * kBlock;
        if (k < nel * kGeo) vg[r] = a.geo[e0 * kGeo + k];
      }
      if (own) {
#pragma unroll
        for (int c = 0; c < NF; ++c) s_q[c][tid] = v[c];
      }
#pragma unroll
      for (int r = 0; r < RG; ++r) {
        const int k = tid + r * kBlock;
        if (k < nel * kGeo) s_geo[k] = vg[r];
      }
    }
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
        const double inv_rho = rcp_nr(rho);
        const double m[3] = {s_q[2][node], s_q[3][node], s_q[4][node]};
        const double u[3] = {m[0] * inv_rho, m[1] * inv_rho, m[2] * inv_rho};
        const double p = (gamma - 1.0) * (e - 0.5 * (m[0] * u[0] + m[1] * u[1] + m[2] * u[2]));
        const double h = (e + p) * inv_rho;
        // flux tensor F[c][x]
        double f[5][3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          f[0][x] = m[x];
          f[1][x] = m[x] * h;
#pragma unroll
          for (int y = 0; y < 3; ++y) f[2 + y][x] = m[y] * u[x] + (x == y ? p : 0.0);
        }
#pragma unroll
        for (int c = 0; c < 5; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r)
            s_w[c * 3 + r][node] = g[r] * f[c][0] + g[3 + r] * f[c][1] + g[6 + r] * f[c][2];
      }
    }
    // face fluxes, scaled by Fscale (exterior values from this thread's slots)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = tid + r * kBlock;
      if (k >= nel * NFN) continue;
      const int fe = k / NFN;
      const int fn = k - fe * NFN;
      const int face = fn / NFP;
      const double* g = &s_geo[fe * kGeo + 9 + 4 * face];
      const double n[3] = {g[0], g[1], g[2]};
      double qm[NF], qp[NF], out[NF];
      const int vn = fe * NP + s_fmask[fn];
#pragma unroll
      for (int c = 0; c < NF; ++c) qm[c] = s_q[c][vn];
      if (pf_ext[r] >= 0) {
#pragma unroll
        for (int c = 0; c < NF; ++c) qp[c] = pf_q[r][c];
      } else {
        Flux::boundary(qm, n, qp);
      }
      Flux::flux(qm, qp, n, c0, out);
#pragma unroll
      for (int c = 0; c < NF; ++c) s_fl[c][fe * NFN + fn] = g[3] * out[c];
    }
    if (has_next) {
      issue_map(sched_at(a.sched.order, pos + p_step));
      issue_gathers();
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
          const double d0 = ROWS ? dr[0][ROWS ? j : 0] : s_dt[j * NP + i];
          const double d1 = ROWS ? dr[1][ROWS ? j : 0] : s_dt[NP * NP + j * NP + i];
          const double d2 = ROWS ? dr[2][ROWS ? j : 0] : s_dt[2 * NP * NP + j * NP + i];
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
#pragma unroll UNROLL_J
        for (int j = 0; j < NP; ++j) {
          const double d0 = ROWS ? dr[0][ROWS ? j : 0] : s_dt[j * NP + i];
          const double d1 = ROWS ? dr[1][ROWS ? j : 0] : s_dt[NP * NP + j * NP + i];
          const double d2 = ROWS ? dr[2][ROWS ? j : 0] : s_dt[2 * NP * NP + j * NP + i];
#pragma unroll
          for (int c = 0; c < 5; ++c)
            acc[c] += d0 * s_w[3 * c][base + j] + d1 * s_w[3 * c + 1][base + j] +
                      d2 * s_w[3 * c + 2][base + j];
        }
      } else {
        double t[6][3];
#pragma unroll
        for (int c = 0; c < 6; ++c) t[c][0] = t[c][1] = t[c][2] = 0.0;
#pragma unroll UNROLL_J
        for (int j = 0; j < NP; ++j) {
          const double d0 = ROWS ? dr[0][ROWS ? j : 0] : s_dt[j * NP + i];
          const double d1 = ROWS ? dr[1][ROWS ? j : 0] : s_dt[NP * NP + j * NP + i];
          const double d2 = ROWS ? dr[2][ROWS ? j : 0] : s_dt[2 * NP * NP + j * NP + i];
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
      for (int c = 0; c < NF; ++c) __builtin_nontemporal_store(acc[c], a.rhs.p[c] + out);
    }
  }
}

// }}}


}  // namespace dg
