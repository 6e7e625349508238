// Euler adjoint (VJP) kernel on the matrix cores, FP64, tetrahedra.
//
// The same transposition as rhs_vjp_kernel (rhs_adjoint.h): with
// mu = M^{-1} lam from the mass pass,
//   volume:  T_{c,r} = A_r mu_c,   F_bar_{c,x} = sum_r rx[x][r] T_{c,r},
//            q_bar += sum_x (dF_x/dq)^T F_bar_x             at every node
//            (here as sum_r (d (F . n_r)/dq)^T T_r, below);
//   faces:   psi(f, a)  = fs_e(f)  sum_b Mf_f[b][a] mu_e (fmask_f[b]),
//            psi'(f, a) = fs_e'(f) sum_b Mf_f[b][a] mu_e'(ext(f, b)),
//            q_bar(fmask_f[a]) += flux VJP at face node a (EulerAdj), which
//            for the conservative flux takes psi - psi' at interior faces.
// Here they run on 16-element tiles, one wave per tile and no block-wide
// barrier after the operators are in LDS, like the forward MFMA kernel
// (rhs_mfma.h):
//   - the contractions are v_mfma_f64_16x16x4f64 (an output block of at most
//     4 rows on 4x4x4);
//   - face f's psi - psi' is one accumulator: P_f (fs mu_e) + Mf_f (-fs' mu_e'),
//     where P_f[a][j] = Mf_f[b][a] at j = fmask_f[b] restricts Mf_f to the
//     volume nodes, so its B operands are the volume phase's mu registers
//     (element n's fs scales lane n's operand).  Where those tiles would not
//     fit (P4) the own traces are loaded and go through Mf_f as well;
//   - face f's result comes out in the C layout, lane (g, n) holding face
//     nodes a = g + 4 v of element n, which is also the B layout of the
//     k-steps v of a contraction over face nodes.  So the lane that gathers
//     the neighbour's mu and q at face node a feeds mu to the contraction
//     and evaluates the flux VJP of that node with its own accumulator;
//   - the tile's state is staged in LDS by contiguous loads where it fits
//     (P1-P3), and the volume term and the own traces read it there;
//   - q_bar of the tile is summed in a per-wave LDS tile [field][n * SP + j]:
//     the volume term is written first, then each face adds its nodes' flux
//     VJPs (a face's nodes are distinct volume nodes, and the faces follow
//     one another in the wave's program order); contiguous stores at the end.
// Tables of faces are face-major: the contraction index of face f at k-step
// ks, lane group g, is face node b = g + 4 ks (0-padded past Nfp).
#pragma once

#include "rhs_adjoint.h"
#include "rhs_mfma.h"

namespace dg {

struct VjpMfmaArgs {
  Schedule sched;                // 16-element tiles
  const double* q[kMaxFields];   // state
  const double* mu[kMaxFields];  // M^{-1} lam per element (mass pass)
  double* qbar[kMaxFields];      // result
  const double* vgeo;            // tile-ordered rx: [t][x][n]
  const double* fgeo;            // tile-ordered face geometry: [t][f * 16 + n][n_x, n_y, n_z, fs]
  const int32_t* ext_v;          // tile-ordered exterior map: [t][lane][f * KSF + v] (-1: none)
  const int32_t* fmask;          // (4 Nfp)
  const double* ext_fscale;      // (K, 4) Fscale of the neighbour across each face
  int32_t face_geo_rx;           // 1: every face satisfies Fscale n = rx^T c_f
  const double* vol_tiles;       // A tiles of A_r: [(r * MB + b) * KS + ks][lane]
  const double* face_tiles;      // A tiles of Mf_f: [f * KSF + ks][lane]
  const double* own_tiles;       // A tiles of P_f: [f * KS + ks][lane]
  int64_t nelem;
  double p0;                     // gamma
};

template <int ORDER>
struct VjpMfmaShape {
  static constexpr int NP = np_of(ORDER);
  static constexpr int NFP = nfp_of(ORDER);
  static constexpr int NFN = kFaces * NFP;
  static constexpr int MB = (NP + 15) / 16;    // 16-row output blocks of A_r mu
  static constexpr int KS = (NP + 3) / 4;      // k-steps over the nodes (j = g * KS + ks)
  static constexpr int KSF = (NFP + 3) / 4;    // k-steps over one face's nodes (b = g + 4 ks)
  static constexpr bool SH = short_last_block(NP);
  static constexpr bool SHF = NFP <= 4;        // face products on 4x4x4
  static constexpr int VOL_TILES = 3 * MB * KS;
  static constexpr int FACE_TILES = kFaces * KSF;
  static constexpr int OWN_TILES = kFaces * KS;
  static constexpr int SP = NP | 1;            // odd row pitch of the LDS tiles
  // the tile's state in LDS and the own psi from the volume mu: both where
  // the LDS holds them (P1-P3; at P4 the block would need 248 KB)
  static constexpr bool STQ = ORDER <= 3;
  static constexpr int EXT = kFaces * KSF;     // exterior-map entries per lane
};

// Whether the exterior map marks neighbours inside the tile, whose state
// the kernel then reads from its staged copy (P2; needs the staged state).
__host__ __device__ constexpr bool vjp_mfma_intile(int order) { return order == 2; }

// Blocks per CU the registers must allow: two (two waves per SIMD) where the
// kernel fits them without spilling (P1, P2), else one.
__host__ __device__ constexpr int vjp_mfma_min_blocks(int order) { return order <= 2 ? 2 : 1; }

template <int ORDER>
__global__ __launch_bounds__(256, vjp_mfma_min_blocks(ORDER)) void rhs_vjp_mfma_kernel(VjpMfmaArgs a) {
  using S = VjpMfmaShape<ORDER>;
  constexpr int NP = S::NP, NFP = S::NFP, MB = S::MB, KS = S::KS, KSF = S::KSF;
  constexpr int SP = S::SP;
  constexpr int NF = 5;
  constexpr int WAVES = 4;
  constexpr bool STQ = S::STQ;

  __shared__ double s_vol[S::VOL_TILES * 64];
  __shared__ double s_face[S::FACE_TILES * 64];
  __shared__ double s_own[STQ ? S::OWN_TILES * 64 : 1];
  __shared__ int s_fmask[S::NFN];
  __shared__ double s_acc[WAVES][NF * 16 * SP];
  __shared__ double s_q[WAVES][STQ ? NF * 16 * SP : 1];

  const int tid = threadIdx.x;
  for (int k = tid; k < S::VOL_TILES * 64; k += 64 * WAVES) s_vol[k] = a.vol_tiles[k];
  for (int k = tid; k < S::FACE_TILES * 64; k += 64 * WAVES) s_face[k] = a.face_tiles[k];
  if constexpr (STQ)
    for (int k = tid; k < S::OWN_TILES * 64; k += 64 * WAVES) s_own[k] = a.own_tiles[k];
  for (int k = tid; k < S::NFN; k += 64 * WAVES) s_fmask[k] = a.fmask[k];
  __syncthreads();

  const int lane = tid & 63;
  const int g = lane >> 4;
  const int n = lane & 15;
  double* const acc = s_acc[tid >> 6];
  double* const sq = s_q[tid >> 6];
  const int slot = block_in_xcd() * WAVES + (tid >> 6);
  const int nslots = blocks_per_xcd() * WAVES;
  const int32_t p_end = a.sched.part[block_xcd() + 1];
  const double gamma = a.p0;

  // neighbour gathers by buffer loads: boundary lanes and padding get an
  // offset past the field's end, which the buffer unit answers with 0 and
  // no memory access (the plan keeps each field below 4 GiB)
  const uint32_t field_bytes = (uint32_t)(a.nelem * NP * 8);
  __amdgpu_buffer_rsrc_t rs_mu[NF], rs_q[NF];
#pragma unroll
  for (int c = 0; c < NF; ++c) {
    rs_mu[c] = __builtin_amdgcn_make_buffer_rsrc((void*)a.mu[c], (short)0, field_bytes, kBufferWord3);
    rs_q[c] = __builtin_amdgcn_make_buffer_rsrc((void*)a.q[c], (short)0, field_bytes, kBufferWord3);
  }
  auto gather = [&](__amdgpu_buffer_rsrc_t rs, uint32_t off) -> double {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
  };

  for (int32_t pos = __builtin_amdgcn_readfirstlane(a.sched.part[block_xcd()] + slot); pos < p_end;
       pos += nslots) {
    const int64_t tile = sched_at(a.sched.order, pos);
    const int64_t gel = tile * 16 + n;
    const int64_t el = gel < a.nelem ? gel : a.nelem - 1;   // clamped for loads
    const int nvalid = (int)(a.nelem - tile * 16 < 16 ? a.nelem - tile * 16 : 16);

    // every face's exterior-map entries (contiguous per lane), the tile's
    // state into LDS (contiguous loads), rx and the volume B operands
    int32_t ext[S::EXT];
    {
      const int32_t* src = a.ext_v + (tile * 64 + lane) * S::EXT;
#pragma unroll
      for (int k = 0; k < S::EXT; ++k) ext[k] = src[k];
    }
    constexpr int NSLOT = (16 * NP + 63) / 64;
    if constexpr (STQ) {
      double v[NF][NSLOT];
#pragma unroll
      for (int c = 0; c < NF; ++c)
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
          const int idx = lane + 64 * s;
          // past the last element: a state with positive pressure, so that
          // the unused columns stay finite
          v[c][s] = idx < nvalid * NP ? a.q[c][tile * 16 * NP + idx] : (c == 0 ? 1.0 : c == 1 ? 2.5 : 0.0);
        }
#pragma unroll
      for (int c = 0; c < NF; ++c)
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
          const int idx = lane + 64 * s;
          if (idx < 16 * NP) sq[c * 16 * SP + (idx / NP) * SP + idx % NP] = v[c][s];
        }
    }
    double rx[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) rx[k] = a.vgeo[(tile * 9 + k) * 16 + n];
    double mv[NF][KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int c = 0; c < NF; ++c) mv[c][ks] = g * KS + ks < NP ? a.mu[c][el * NP + g * KS + ks] : 0.0;
    if constexpr (STQ) wave_lds_fence();
    // state of element n at node j
    auto qat = [&](int c, int j) -> double {
      if constexpr (STQ) return sq[c * 16 * SP + n * SP + j];
      else return a.q[c][el * NP + j];
    };

    // {{{ volume: (d (F . n) / dq)^T is linear in n, so
    // sum_x (dF_x/dq)^T sum_r rx[x][r] T_r = sum_r (d (F . n_r) / dq)^T T_r
    // with n_r = (rx[0][r], rx[1][r], rx[2][r]): one direction r at a time,
    // 5 accumulators live instead of 15

#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const bool sb = S::SH && b == MB - 1;
      double qn[4][NF], out[4][NF];
#pragma unroll
      for (int v = 0; v < ncomp(sb); ++v) {
        const int j = 16 * b + g + 4 * v;
#pragma unroll
        for (int c = 0; c < NF; ++c) {
          qn[v][c] = j < NP ? qat(c, j) : 1.0;
          out[v][c] = 0.0;
        }
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        f64x4 t[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) t[c] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const double av = s_vol[((r * MB + b) * KS + ks) * 64 + lane];
#pragma unroll
          for (int c = 0; c < NF; ++c) mma(t[c], av, mv[c][ks], sb);
        }
        const double nr[3] = {rx[r], rx[3 + r], rx[6 + r]};
#pragma unroll
        for (int v = 0; v < ncomp(sb); ++v) {
          if (16 * b + g + 4 * v < NP) {
            double tv[NF];
#pragma unroll
            for (int c = 0; c < NF; ++c) tv[c] = t[c][v];
            EulerAdj::nflux_vjp(qn[v], nr, gamma, tv, out[v]);
          }
        }
      }
#pragma unroll
      for (int v = 0; v < ncomp(sb); ++v) {
        const int j = 16 * b + g + 4 * v;
        if (j < NP) {
#pragma unroll
          for (int c = 0; c < NF; ++c) acc[c * 16 * SP + n * SP + j] = out[v][c];
        }
      }
    }

    // }}}

    // {{{ faces: lane (g, n) owns face nodes a = g + 4 v of element n

#pragma unroll 1
    for (int f = 0; f < kFaces; ++f) {
      // face f of element n: n and Fscale formed from rx (Fscale n = rx^T
      // c_f) where the plan found that on every face, else from the table
      double nrm[3], fs;
      if (a.face_geo_rx) {
        double m[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          const double c0 = f == 3 ? -1.0 : (f == 2 ? 1.0 : 0.0);
          const double c1 = f == 1 ? -1.0 : (f == 2 ? 1.0 : 0.0);
          const double c2 = f == 0 ? -1.0 : (f == 2 ? 1.0 : 0.0);
          m[x] = c0 * rx[3 * x] + c1 * rx[3 * x + 1] + c2 * rx[3 * x + 2];
        }
        fs = sqrt_nr(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        const double inv = rcp_nr(fs);
#pragma unroll
        for (int x = 0; x < 3; ++x) nrm[x] = m[x] * inv;
      } else {
        const double2* fg = reinterpret_cast<const double2*>(a.fgeo) + (tile * 64 + f * 16 + n) * 2;
        const double2 g01 = fg[0], g23 = fg[1];
        nrm[0] = g01.x;
        nrm[1] = g01.y;
        nrm[2] = g23.x;
        fs = g23.y;
      }
      const double xfs = a.ext_fscale[el * kFaces + f];
      // the face's exterior entries (a runtime face index into ext[] would
      // put it in scratch memory)
      int32_t fe[KSF];
#pragma unroll
      for (int v = 0; v < KSF; ++v) {
        fe[v] = ext[v];
#pragma unroll
        for (int f2 = 1; f2 < kFaces; ++f2)
          if (f == f2) fe[v] = ext[f2 * KSF + v];
      }
      // Neighbours inside the tile (IT: kInTile | local element << 8 | node)
      // take q from the staged tile, the others by buffer loads whose
      // out-of-range lanes (those, and boundary nodes) touch no memory; mu
      // of an in-tile neighbour is loaded at its global index.  Measured on
      // MI355X: P2 -1.2%, P1 +1.4%, P3 +28%, so at P2 only.
      constexpr bool IT = vjp_mfma_intile(ORDER);
      double mx[NF][KSF], qx[NF][KSF], mo[STQ ? 1 : NF][KSF];
#pragma unroll
      for (int v = 0; v < KSF; ++v) {
        const int32_t e = fe[v];
        const bool loc = IT && e >= kInTile;
        const bool glob = e >= 0 && !loc;
        const int lel = (e >> 8) & 15, lnd = e & 255;
        const uint32_t goff = glob ? (uint32_t)e * 8u : field_bytes;
        const uint32_t moff =
            loc ? (uint32_t)(((int64_t)tile * 16 + lel) * NP + lnd) * 8u : goff;
        const bool ok = g + 4 * v < NFP;
        const int nd = ok ? s_fmask[f * NFP + g + 4 * v] : 0;
#pragma unroll
        for (int c = 0; c < NF; ++c) {
          mx[c][v] = gather(rs_mu[c], moff);
          const double gv = gather(rs_q[c], goff);
          if constexpr (IT) {
            const double lv = sq[c * 16 * SP + (loc ? lel * SP + lnd : 0)];
            qx[c][v] = loc ? lv : gv;
          } else {
            qx[c][v] = gv;
          }
          if constexpr (!STQ) mo[c][v] = ok ? a.mu[c][el * NP + nd] : 0.0;
        }
      }
      // psi - psi' in one accumulator: the B operands carry fs and -fs'
      // (element n's factors); at a boundary face mx is 0 and it is psi
      f64x4 pd[NF];
#pragma unroll
      for (int c = 0; c < NF; ++c) pd[c] = f64x4{0.0, 0.0, 0.0, 0.0};
      if constexpr (STQ) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const double av = s_own[(f * KS + ks) * 64 + lane];
#pragma unroll
          for (int c = 0; c < NF; ++c) mma(pd[c], av, fs * mv[c][ks], S::SHF);
        }
      }
#pragma unroll
      for (int ks = 0; ks < KSF; ++ks) {
        const double av = s_face[(f * KSF + ks) * 64 + lane];
#pragma unroll
        for (int c = 0; c < NF; ++c) {
          if constexpr (!STQ) mma(pd[c], av, fs * mo[c][ks], S::SHF);
          mma(pd[c], av, -xfs * mx[c][ks], S::SHF);
        }
      }
#pragma unroll
      for (int v = 0; v < (S::SHF ? 1 : KSF); ++v) {
        if (g + 4 * v < NFP) {
          const int nd = s_fmask[f * NFP + g + 4 * v];
          double qm[NF], qp[NF], gm[NF];
#pragma unroll
          for (int c = 0; c < NF; ++c) {
            qm[c] = qat(c, nd);
            qp[c] = qx[c][v];
          }
          if (fe[v] < 0) {
            double psi[NF], gp[NF];
#pragma unroll
            for (int c = 0; c < NF; ++c) psi[c] = pd[c][v];
            EulerFlux::boundary(qm, nrm, qp);
            EulerAdj::vjp(qm, qp, nrm, gamma, psi, gm, gp);
            EulerAdj::bcT(nrm, gp);
#pragma unroll
            for (int c = 0; c < NF; ++c) gm[c] += gp[c];
          } else {
            double pdv[NF];
#pragma unroll
            for (int c = 0; c < NF; ++c) pdv[c] = pd[c][v];
            EulerAdj::vjp_m(qm, qp, nrm, gamma, pdv, gm);
          }
#pragma unroll
          for (int c = 0; c < NF; ++c) acc[c * 16 * SP + n * SP + nd] += gm[c];
        }
      }
    }

    // }}}

    // {{{ output: contiguous non-temporal stores of the LDS tile

    wave_lds_fence();
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      double* dst = a.qbar[c] + tile * 16 * NP;
      double vals[NSLOT];
#pragma unroll
      for (int s = 0; s < NSLOT; ++s) {
        const int idx = lane + 64 * s;
        vals[s] = idx < 16 * NP ? acc[c * 16 * SP + (idx / NP) * SP + idx % NP] : 0.0;
      }
#pragma unroll
      for (int s = 0; s < NSLOT; ++s) {
        const int idx = lane + 64 * s;
        if (idx < nvalid * NP) __builtin_nontemporal_store(vals[s], dst + idx);
      }
    }
    wave_lds_fence();

    // }}}
  }
}

}  // namespace dg
