/* CPU restatement of the fused DG right-hand sides, OpenMP over elements
 * (TEST INFRASTRUCTURE ONLY: the CPU baseline bench.py times, and a second
 * checker of the fused operator forms; the engine never links it).
 *
 * It evaluates, per affine tetrahedron e, the operators in the reference
 * optimizer's fused forms (arraycontext/mass_inverse_fuser.py:20-74):
 *   weak divergence  xre,rij,xej->ei  with D = M^{-1} S_r^T (wave, Euler)
 *   strong gradient  re,rij,ej->ei    with D = D_r (Maxwell's curl)
 *   LIFT             ifj,fe,fej->ei   with LIFT = M^{-1} E, Fscale = sJ/J
 * and the numerical fluxes of
 *   wave     equations/wave.py:80-200 (upwind, mirror walls u+ = -u-, v+ = v-)
 *   Euler    equations/euler.py:168-174 (Lax-Friedrichs, lambda = max(|u| + c)
 *            over both traces, reflecting walls)
 *   Maxwell  equations/maxwell.py:57-59 (flux_type alpha, PEC walls)
 * on the data layout the HIP plan takes: fields (K, Np) row-major, geometry
 * (K, 25) = rx[x][r] (9) then per face (n_x, n_y, n_z, Fscale), exterior map
 * (K, 4 Nfp) with boundary entries -(own)-1.
 *
 * Build: gcc -O3 -march=x86-64-v3 -fopenmp -shared -fPIC (oracle/Makefile). */

#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NFACES 4
#define MAXNP 35
#define MAXNF 6
#define MAXNFN (NFACES * 15)

static void wave_flux(const double* qm, const double* qp, const double* n, double c, double* f) {
  const double uavg = 0.5 * (qm[0] + qp[0]), ujump = qp[0] - qm[0];
  double vn_avg = 0.0, vn_jump = 0.0;
  for (int x = 0; x < 3; ++x) {
    vn_avg += 0.5 * (qm[1 + x] + qp[1 + x]) * n[x];
    vn_jump += (qp[1 + x] - qm[1 + x]) * n[x];
  }
  f[0] = c * (vn_avg + 0.5 * ujump);
  for (int x = 0; x < 3; ++x) f[1 + x] = c * (uavg * n[x] + 0.5 * vn_jump * n[x]);
}

static void euler_tensor(const double* q, double gamma, double F[5][3], double* speed) {
  const double rho = q[0], e = q[1];
  const double u[3] = {q[2] / rho, q[3] / rho, q[4] / rho};
  const double p = (gamma - 1.0) * (e - 0.5 * (q[2] * u[0] + q[3] * u[1] + q[4] * u[2]));
  for (int x = 0; x < 3; ++x) {
    F[0][x] = q[2 + x];
    F[1][x] = (e + p) * u[x];
    for (int i = 0; i < 3; ++i) F[2 + i][x] = q[2 + i] * u[x] + (i == x ? p : 0.0);
  }
  if (speed) *speed = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) + sqrt(gamma * (p / rho));
}

static void euler_flux(const double* qm, const double* qp, const double* n, double gamma, double* f) {
  double Fm[5][3], Fp[5][3], sm, sp;
  euler_tensor(qm, gamma, Fm, &sm);
  euler_tensor(qp, gamma, Fp, &sp);
  const double lam = sm > sp ? sm : sp;
  for (int c = 0; c < 5; ++c) {
    double avg = 0.0;
    for (int x = 0; x < 3; ++x) avg += 0.5 * (Fm[c][x] + Fp[c][x]) * n[x];
    f[c] = avg + 0.5 * lam * (qm[c] - qp[c]);
  }
}

static void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

static void maxwell_flux(const double* qm, const double* qp, const double* n, double alpha, double* f) {
  double de[3], dh[3], nde[3], ndh[3], nnde[3], nndh[3];
  for (int i = 0; i < 3; ++i) {
    de[i] = qp[i] - qm[i];
    dh[i] = qp[3 + i] - qm[3 + i];
  }
  cross(n, de, nde);
  cross(n, dh, ndh);
  cross(n, nde, nnde);
  cross(n, ndh, nndh);
  for (int i = 0; i < 3; ++i) {
    f[i] = -0.5 * ndh[i] + 0.5 * alpha * nnde[i];
    f[3 + i] = 0.5 * nde[i] + 0.5 * alpha * nndh[i];
  }
}

static void boundary_state(int eq, int nf, const double* qm, const double* n, double* qb) {
  for (int c = 0; c < nf; ++c) qb[c] = qm[c];
  if (eq == 0) {
    qb[0] = -qm[0];
  } else if (eq == 1) {
    const double mn = qm[2] * n[0] + qm[3] * n[1] + qm[4] * n[2];
    for (int x = 0; x < 3; ++x) qb[2 + x] = qm[2 + x] - 2.0 * mn * n[x];
  } else {
    for (int i = 0; i < 3; ++i) qb[i] = -qm[i];
  }
}

static int levi(int o, int x, int a) {
  if (o == x || x == a || o == a) return 0;
  return ((x - o + 3) % 3 == 1) ? 1 : -1;
}

/* equation: 0 wave (param c), 1 Euler (gamma), 2 Maxwell (alpha).
 * d_weak, d_strong: (3, Np, Np); lift: (Np, 4 Nfp); fmask: (4, Nfp).
 * Returns 0, or 1 on a bad argument. */
int dg_cpu_rhs(int equation, int np, int nfp, int64_t nelem, const double* d_weak,
               const double* d_strong, const double* lift, const int32_t* fmask,
               const double* geo, const int32_t* emap, double param,
               const double* const* q, double* const* rhs) {
  const int nf = equation == 0 ? 4 : equation == 1 ? 5 : equation == 2 ? 6 : -1;
  if (nf < 0 || np > MAXNP || np < 1 || nfp > 15 || nelem < 0) return 1;
  const int nfn = NFACES * nfp;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nelem; ++e) {
    const double* g = geo + e * 25;
    double out[MAXNF][MAXNP];
    double qe[MAXNF][MAXNP];
    for (int c = 0; c < nf; ++c)
      for (int j = 0; j < np; ++j) {
        qe[c][j] = q[c][e * np + j];
        out[c][j] = 0.0;
      }
    /* {{{ volume */
    if (equation == 2) {
      /* strong derivatives d_x F_a = sum_r rx[x][r] D_r F_a */
      for (int a = 0; a < 6; ++a) {
        double dr[3][MAXNP];
        for (int r = 0; r < 3; ++r)
          for (int i = 0; i < np; ++i) {
            double s = 0.0;
            const double* row = d_strong + ((size_t)r * np + i) * np;
            for (int j = 0; j < np; ++j) s += row[j] * qe[a][j];
            dr[r][i] = s;
          }
        const int comp = a % 3, is_h = a >= 3;
        for (int x = 0; x < 3; ++x)
          for (int o = 0; o < 3; ++o) {
            const int s = levi(o, x, comp);
            if (!s) continue;
            const int dst = is_h ? o : 3 + o;
            const double sg = is_h ? (double)s : -(double)s;
            for (int i = 0; i < np; ++i)
              out[dst][i] += sg * (g[3 * x] * dr[0][i] + g[3 * x + 1] * dr[1][i] +
                                   g[3 * x + 2] * dr[2][i]);
          }
      }
    } else {
      /* weak form: rhs_c += sum_r D_r w_{c,r}, w_{c,r} = sum_x rx[x][r] W_cx(q) */
      double w[MAXNF][3][MAXNP];
      for (int j = 0; j < np; ++j) {
        double W[5][3];
        if (equation == 0) {
          for (int x = 0; x < 3; ++x) {
            W[0][x] = -param * qe[1 + x][j];
            for (int y = 0; y < 3; ++y) W[1 + y][x] = (x == y) ? -param * qe[0][j] : 0.0;
          }
        } else {
          double qq[5];
          for (int c = 0; c < 5; ++c) qq[c] = qe[c][j];
          euler_tensor(qq, param, W, NULL);
        }
        for (int c = 0; c < nf; ++c)
          for (int r = 0; r < 3; ++r)
            w[c][r][j] = g[r] * W[c][0] + g[3 + r] * W[c][1] + g[6 + r] * W[c][2];
      }
      for (int c = 0; c < nf; ++c)
        for (int r = 0; r < 3; ++r)
          for (int i = 0; i < np; ++i) {
            double s = 0.0;
            const double* row = d_weak + ((size_t)r * np + i) * np;
            for (int j = 0; j < np; ++j) s += row[j] * w[c][r][j];
            out[c][i] += s;
          }
    }
    /* }}} */
    /* {{{ surface */
    double flux[MAXNF][MAXNFN];
    for (int f = 0; f < NFACES; ++f) {
      const double* gf = g + 9 + 4 * f;
      const double n[3] = {gf[0], gf[1], gf[2]};
      for (int j = 0; j < nfp; ++j) {
        const int fn = f * nfp + j;
        const int vn = fmask[fn];
        double qm[MAXNF], qp[MAXNF], fl[MAXNF];
        for (int c = 0; c < nf; ++c) qm[c] = qe[c][vn];
        const int32_t ext = emap[e * nfn + fn];
        if (ext >= 0)
          for (int c = 0; c < nf; ++c) qp[c] = q[c][ext];
        else
          boundary_state(equation, nf, qm, n, qp);
        if (equation == 0)
          wave_flux(qm, qp, n, param, fl);
        else if (equation == 1)
          euler_flux(qm, qp, n, param, fl);
        else
          maxwell_flux(qm, qp, n, param, fl);
        const double sign = equation == 0 ? 1.0 : -1.0;
        for (int c = 0; c < nf; ++c) flux[c][fn] = sign * gf[3] * fl[c];
      }
    }
    for (int c = 0; c < nf; ++c)
      for (int i = 0; i < np; ++i) {
        double s = 0.0;
        const double* row = lift + (size_t)i * nfn;
        for (int k = 0; k < nfn; ++k) s += row[k] * flux[c][k];
        out[c][i] += s;
      }
    /* }}} */
    for (int c = 0; c < nf; ++c)
      for (int i = 0; i < np; ++i) rhs[c][e * np + i] = out[c][i];
  }
  return 0;
}

/* Threads of the parallel element loop (OpenMP). */
void dg_cpu_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
