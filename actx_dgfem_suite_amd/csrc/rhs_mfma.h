// Fused FP64 RHS kernel on the matrix cores (v_mfma_f64_16x16x4f64).
//
// A tile is 16 consecutive elements (the N dimension of the MFMA).  Each wave
// takes tiles from its XCD's slice of the plan's locality schedule and works
// on them independently of the other waves of its block -- the only
// block-wide barrier is the initial load of the constant operators into LDS
// -- so the memory, LDS, VALU and MFMA phases of different waves overlap
// freely.
//
// Per tile:
//   volume: B operands straight from HBM into registers (lane group g loads a
//     contiguous run of each element's nodes), A operands (pre-tiled
//     operators) from LDS, rx from the tile-ordered geometry (one cache line
//     per load);
//   surface, C LIFT k-steps at a time: the wave evaluates the numerical flux
//     in a gather-friendly order (consecutive lanes = consecutive nodes of one
//     face, so the exterior gathers of neighbouring lanes hit the same
//     neighbour element), with normals, Fscale and exterior-map entries read
//     from per-tile tables laid out in that item order (contiguous loads),
//     hands it to itself through a small LDS buffer in MFMA B order and runs
//     those LIFT k-steps;
//   output: accumulators go through the same buffer to contiguous stores.
//
// Operand layouts of v_mfma_f64_16x16x4f64 (lane l, g = l / 16, n = l % 16),
// measured on gfx950:
//   A (16 x 4):  lane l holds A[m = n][k = g]
//   B (4 x 16):  lane l holds B[k = g][col = n]
//   C (16 x 16): lane l, register v holds C[row = g + 4 v][col = n]
// (the f64 output interleaves rows across lane groups, unlike the f32 16x16x4
// form whose rows are 4 g + v).  The contraction index k = g at k-step ks
// stands for node j = g * KS + ks.  For LIFT the same rule maps lane group g
// to face g (face nodes are face-major).  The host writes A tiles in the same
// order:
//   tile[(r * MB + b) * KS + ks][l] = D_r[16 b + n][g * KS + ks]   (0-padded)
//
// Element n of the tile is fixed per lane, so per-element geometric factors
// are applied in registers to MFMA inputs and outputs:
//   gradient form  t_r = D_r u,           out_x = sum_r rx[x][r] t_r
//   divergence     out = sum_r D_r w_r,   w_r = sum_x rx[x][r] F_x
// The two share every A tile.
#pragma once

#include "dg_common.h"

namespace dg {

typedef double f64x4 __attribute__((ext_vector_type(4)));


__device__ inline f64x4 mfma_f64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Output blocks of at most 4 rows (the last block of P1, P3 and P4: Np = 4,
// 20 = 16 + 4, 35 = 32 + 3) run on v_mfma_f64_4x4x4f64, which takes 0.28x the
// time of the 16x16x4 form on gfx950.  Its four 4x4 blocks tile the 16
// columns, and, measured on gfx950, its B operand layout is the 16x16x4 one
// (lane l: B[k = l / 16][col = l % 16]) and its result is the 16x16x4
// result's register v = 0 (lane l: C[row = l / 16][col = l % 16]); only A
// differs (lane l: A[row = l % 4][k = l / 16]), which the host's A tiles
// follow.  So a short block keeps the same B operands and uses component 0 of
// its accumulator.
__host__ __device__ constexpr bool short_last_block(int np) {
  return np - 16 * ((np + 15) / 16 - 1) <= 4;
}

// acc += A B on output block b (short: component 0 only).
__device__ inline void mma(f64x4& acc, double a, double b, bool short_block) {
  if (short_block)
    acc[0] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc[0], 0, 0, 0);
  else
    acc = mfma_f64(a, b, acc);
}

// Components of an accumulator that hold rows: 1 for a short block, else 4.
__device__ constexpr int ncomp(bool short_block) { return short_block ? 1 : 4; }

template <int ORDER>
struct MfmaShape {
  static constexpr int NP = np_of(ORDER);
  static constexpr int NFP = nfp_of(ORDER);
  static constexpr int NFN = kFaces * NFP;
  static constexpr int MB = (NP + 15) / 16;   // 16-row blocks of the output
  static constexpr int KS = (NP + 3) / 4;     // k-steps of one D_r contraction
  static constexpr int KSL = NFP;             // k-steps of the LIFT contraction
  static constexpr int OP_TILES = 3 * MB * KS;
  static constexpr int LIFT_TILES = MB * KSL;
};

// Whether a wave loads the exterior-map entries of its next flux chunk ahead
// of time.  It costs C registers per lane; measured on MI355X, it pays where
// it leaves the occupancy unchanged (Euler P2 -2%, wave P4 -6%) and does not
// for wave P1 and P3, where those registers cost a wave per SIMD (wave P3 +11%).
template <int EQ, int ORDER>
constexpr bool prefetch_ext() {
  return !(EQ == DG_EQ_WAVE && (ORDER == 1 || ORDER == 3));
}

// Waves per block.  Measured on MI355X at P4, blocks of 8 waves (two waves
// per SIMD, the 64 KB of operator tiles shared by twice the waves, flux chunks
// of 3 or 4 k-steps to fit) are slower than 4 waves at one wave per SIMD:
// wave P4 on 584k tets 949-1042 against 879 us.  Twice the tiles in flight
// hold about 4.6 MB of state per XCD, more than its 4 MB L2, which the flux
// producer's own-trace and neighbour reads rely on.
__host__ __device__ constexpr int mfma_waves_of(int eq, int order) { return 4; }

// LIFT k-steps whose flux one wave stages in LDS at a time (P1, P2: one
// chunk; P3, P4: 5 nodes; 3 at P4 measured 5% slower on wave P4).  Host code
// sizes the flux-item tables with it.
__host__ __device__ constexpr int flux_chunk_of(int eq, int order) {
  return nfp_of(order) <= 6 ? nfp_of(order) : 5;
}
template <int EQ, int ORDER>
constexpr int flux_chunk() { return flux_chunk_of(EQ, ORDER); }

// Flux items of one tile: chunk ch, slot it, lane l handle item it*64 + l of
// that chunk, i.e. face-pair p = item / C (p = face * 16 + element) at LIFT
// k-step ch * C + item % C.  ext_t holds the exterior-map entry of every
// item in this order, so each slot's entries are one contiguous 256-byte load.
__host__ __device__ constexpr int flux_items_of(int eq, int order) {
  return ((nfp_of(order) + flux_chunk_of(eq, order) - 1) / flux_chunk_of(eq, order)) *
         flux_chunk_of(eq, order) * 64;
}

// Whether the surface phase evaluates the flux directly in the LIFT B-operand
// layout: at k-step ks, lane (g, n) evaluates the flux at one face node of
// element n (the face-node pair of lane group g, direct_pair_of below), so
// each lane forms the flux value its own B operand needs, the LIFT A tiles
// follow that k order and no flux goes through LDS.  The exterior map comes
// in the same order (ext_d).  One gather instruction then reads 4 face nodes
// of each of 16 neighbours.  It frees the flux buffer's LDS, which pays for
// the staged own traces at wave P3 and for a third wave per SIMD at Maxwell
// P2, and it is what lets the wave operator LIFT its velocity flux in
// reference normals (ref_normal_lift).  Measured on MI355X against the
// chunked producer: Maxwell P2 79 against 88 us for the LDS-staged kernel,
// Maxwell P4 -3%; with the reference-normal LIFT, wave P2 -1.5%, P4 -2.3%;
// Euler P4 (packed k-steps) -6%; slower elsewhere (Euler P2 +17%, P3 +22%,
// Maxwell P3 +26%, wave P1 +15%).
__host__ __device__ constexpr bool direct_flux_of(int eq, int order) {
  return (eq == DG_EQ_WAVE && order >= 2) || (eq == DG_EQ_MAXWELL && (order == 2 || order == 4)) ||
         (eq == DG_EQ_EULER && order == 4);
}
template <int EQ, int ORDER>
constexpr bool direct_flux() { return direct_flux_of(EQ, ORDER); }

// The direct flux's k order.  Padded: k-step ks covers face ks / KPF, lane
// group g its node 4 (ks % KPF) + g (KPF = ceil(Nfp / 4) k-steps per face,
// nodes past Nfp zero-padded).  Packed: k-step ks covers the face-node pair
// p = 4 ks + g, i.e. face p / Nfp and node p % Nfp, so the 4 Nfp pairs fill
// Nfp k-steps with no padding and a k-step may straddle two faces.  Packed
// except where the straddling k-steps would cost the wave operator's
// reference-normal LIFT (below) more MFMAs than the padding: wave P1, P4.
__host__ __device__ constexpr bool packed_direct_of(int eq, int order) {
  return eq != DG_EQ_WAVE || order == 2 || order == 3;
}
__host__ __device__ constexpr int direct_ksteps_of(int eq, int order) {
  return packed_direct_of(eq, order) ? nfp_of(order) : kFaces * ((nfp_of(order) + 3) / 4);
}
// Face-node pair (face * Nfp + node) of k-step ks, lane group g, or -1 for a
// padded slot.
__host__ __device__ constexpr int direct_pair_of(int eq, int order, int ks, int g) {
  return packed_direct_of(eq, order)
             ? 4 * ks + g
             : (4 * (ks % ((nfp_of(order) + 3) / 4)) + g < nfp_of(order)
                    ? (ks / ((nfp_of(order) + 3) / 4)) * nfp_of(order) +
                          4 * (ks % ((nfp_of(order) + 3) / 4)) + g
                    : -1);
}

// Whether the wave operator LIFTs its velocity flux in reference normals.
// The upwind flux's velocity part is s n (s scalar per face node), and on an
// affine tetrahedron Fscale_f n_f = rx^T c_f (Nanson's formula), where
// c_f = the reference face's outward normal times its area ratio is
// (0, 0, -1), (0, -1, 0), (1, 1, 1), (-1, 0, 0) for faces 0..3 (the plan
// checks every element's geometry against it).  So
//   LIFT(Fscale s n_x) = sum_r rx[x][r] sum_f c_{f,r} LIFT_f s_f,
// and the velocity flux is LIFTed into the per-r gradient accumulators
// t_r = D_r u before their combination with rx: one MFMA per k-step on an
// axis face and three on the slanted face, against three (one per velocity
// field) before.  With the direct flux layout (one face per lane group, or
// two in a straddling packed k-step) the MFMAs a k-step needs are known at
// compile time.  Measured on MI355X: wave P3 -10%.
__host__ __device__ constexpr bool ref_normal_lift_of(int eq, int order) {
  return eq == DG_EQ_WAVE && direct_flux_of(eq, order);
}
template <int EQ, int ORDER>
constexpr bool ref_normal_lift() { return ref_normal_lift_of(EQ, ORDER); }

// Whether the flux producer takes a boundary state that is the interior trace
// with per-field signs (wave, Maxwell PEC) by negating the loaded own trace
// instead of forming Flux::boundary and selecting per field, and (wave with
// the reference-normal LIFT) evaluates the upwind flux from the two normal
// velocities.  Fewer VALU instructions: measured on MI355X wave P2 -6%, P3
// -3.6%, Maxwell P3 -2%, wave P4 -2.3% (with its staged state; before that
// it measured +2-3% there).
template <int EQ, int ORDER>
constexpr bool lean_flux() {
  return FluxOf<EQ>::type::mirror_sign(0) != 0;
}

// c_{f,r}: reference face normal times area ratio (faces of the reference
// element in its face numbering).
__host__ __device__ constexpr int ref_face_coef(int f, int r) {
  return f == 2 ? 1 : ((f == 0 && r == 2) || (f == 1 && r == 1) || (f == 3 && r == 0) ? -1 : 0);
}

// Whether the volume phase keeps a copy of the tile's state in LDS (as it
// loads the B operands) so that the flux producer reads its own traces from
// there instead of issuing NF more global loads per flux item.  It costs
// NF * 16 * Np doubles of LDS per wave.  Measured on MI355X, it pays where the
// blocks per CU stay what the registers allow (wave P1 -16%, wave P2 -9%,
// Euler P3 -7%, Maxwell P1 -18%, Maxwell P3 -11%, Maxwell P2 even) and not
// where the LDS halves them (wave P3 +8%, Euler P2 +18%).  At P4 it fits the
// wave operator only (146 KB per block with the direct flux): FP64 wave P4
// even, and with it the in-tile neighbours, which the FP32 mixed kernel
// reads from the same map, FP32 wave P4 -9%.
__host__ __device__ constexpr bool stage_own_traces_of(int eq, int order) {
  return direct_flux_of(eq, order) ? (order <= 3 || eq == DG_EQ_WAVE)
                                   : (eq == DG_EQ_WAVE ? order <= 2
                                                       : (eq == DG_EQ_EULER ? order == 3 : order <= 3));
}
template <int EQ, int ORDER>
constexpr bool stage_own_traces() { return stage_own_traces_of(EQ, ORDER); }

// Direct flux: the exterior-map entries of neighbours inside the tile (59% of
// the interior face nodes on the benchmark meshes) point into the staged
// state instead: bit 30 set, the neighbour's place in the tile in bits 8-11
// and its node in bits 0-7.  The other neighbours come by buffer loads with
// 32-bit byte offsets, and the lanes that read the staged state (or a
// boundary node, whose state is its own trace) give an offset past the
// field's end, which the buffer unit answers without a memory access.
// Measured on MI355X: wave P2 -8% to -17%, wave P3 -8% on 384k tets (even on
// 162k tets, whose state stays in the last-level cache).  Needs each field
// below 4 GiB (the plan checks).
__host__ __device__ constexpr bool intile_neighbours_of(int eq, int order) {
  return direct_flux_of(eq, order) && stage_own_traces_of(eq, order);
}

// Whether a wave copies its tile's face geometry (64 face-element pairs x
// (n, Fscale), 2 KB of the tile-ordered table) into LDS with two 16-byte loads
// per lane when it starts the tile, so that the flux producer reads normals
// and Fscale from LDS instead of issuing two global loads per flux item.  On
// where the 2 KB per wave leave the blocks per CU unchanged; measured on
// MI355X: Euler P1-P3 -2% to -5%, wave P4 -3%, wave P3 and Maxwell even.
template <int EQ, int ORDER>
constexpr bool face_geometry_in_lds() {
  if constexpr (EQ == DG_EQ_WAVE) return ORDER >= 3;
  if constexpr (EQ == DG_EQ_EULER) return true;
  return ORDER >= 2;
}

// Whether the flux producer gathers without a branch: a boundary node loads
// its own trace (always a valid address) and takes the boundary state by
// selection, so the C items of a chunk are one basic block and the compiler
// issues their map loads, then all their own-trace and neighbour loads,
// together, instead of waiting out one item's round trip before the next
// item's gathers.  Measured on MI355X, bit-identical results: wave P2-P4 -5%
// to -12%, Euler P2-P4 -4% to -6%, Maxwell P3-P4 -7%; at P1 (3 items per
// chunk) wave +14%, Maxwell +9%, Euler even, and Maxwell P2 +2.5%.
template <int EQ, int ORDER>
constexpr bool branch_free_gathers() {
  return ORDER >= 2 && !(EQ == DG_EQ_MAXWELL && ORDER == 2);
}

// Whether a wave loads the next tile's volume operands (its state at the
// volume B-operand nodes and rx) in the output phase of the current tile,
// before the output's stores: the registers are free there, and loads issued
// ahead of the stores do not wait for them (one counter covers both).
// Measured on MI355X: Euler P3 -3%, wave P2 -2%, Euler P2 -1.4%; elsewhere
// the compiler keeps fewer of the surface phase's loads in flight around the
// held operands (wave P3 +9%, Maxwell P2 +9%, wave and Maxwell P4 +25-30%).
template <int EQ, int ORDER>
constexpr bool pipe_volume() {
  return (EQ == DG_EQ_EULER && (ORDER == 2 || ORDER == 3)) || (EQ == DG_EQ_WAVE && ORDER == 2);
}

constexpr int kFluxRow = 65;                 // padded flux row (spreads LDS banks)

// Row pitch (doubles) of one element's nodes in the per-wave LDS copies of a
// tile (staged state, output tile): odd, so that the 16 elements a lane group
// reads or writes at one node fall in distinct LDS banks (measured: Euler P2
// -2%, wave P3 -1%), unless it would cost a block per CU.
__host__ __device__ constexpr int stage_pitch(int np) { return np | 1; }

// Per-wave LDS buffer (doubles): one flux chunk, or one field of the output
// tile.  With the direct flux and the staged state, the output tile goes
// through the stage (dead after the surface phase) and there is no buffer.
template <int EQ, int ORDER>
constexpr int wave_buffer(int pitch) {
  constexpr int flux = direct_flux<EQ, ORDER>() ? 0 : flux_chunk<EQ, ORDER>() * nfields_of(EQ) * kFluxRow;
  const int out = 16 * pitch;
  if constexpr (direct_flux<EQ, ORDER>() && stage_own_traces<EQ, ORDER>()) return 1;
  return flux > out ? flux : out;
}

// Orders this wave's LDS accesses across lanes: waits for the wave's
// outstanding LDS operations and stops the compiler from moving memory
// accesses across this point.  (A wavefront-scope fence does not order LDS
// accesses of different lanes against each other.)
__device__ inline void wave_lds_fence() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Levi-Civita sign of (o, x, a) for the curl: (curl F)_o = eps_{o x a} d_x F_a.
__device__ constexpr int levi(int o, int x, int a) {
  return (o == x || x == a || o == a) ? 0
         : (((x - o + 3) % 3 == 1) ? 1 : -1);
}

// Unroll depth of the volume k-step loops: enough independent MFMAs and loads
// in flight without hoisting every load of the tile into registers.
// Measured on MI355X against 3: Maxwell P4 -7.9%, wave P4 -7% to -10%, Euler
// P4 -4% (nine k-steps; a full unroll measures the same), P1-P3 even; 1 is
// 5-39% slower.
#define KUNROLL 2
// Unroll depth of the direct-flux k-step loop (measured: 3 against 2 and 4,
// except Maxwell P4, 4).
#define DUNROLL 3
#define DUNROLL_P4 4

// POST: the kernel multiplies each element's RHS by post_sign_c |J| M_ref
// before storing it, one more contraction per field with the output tile as
// the B operand -- the last step of the linear operators' adjoint
// R M L R M^{-1} lam, fused so that it costs no pass over HBM.
// STAGE: instead of storing the RHS, the kernel applies the LSERK4 stage
// res = a res + dt rhs, q_out = beta q + b res to it in its output phase, so
// the RHS makes no round trip through HBM between the two (beta = 1 forward;
// the reverse stage in the adjoint variables is res = q_hat, a = 1).
template <int EQ, int ORDER, bool POST = false, bool STAGE = false>
__global__ __launch_bounds__(64 * mfma_waves_of(EQ, ORDER)) void rhs_mfma_kernel(RhsArgs a) {
  using S = MfmaShape<ORDER>;
  using Flux = typename FluxOf<EQ>::type;
  constexpr int NP = S::NP, NFP = S::NFP, NFN = S::NFN, MB = S::MB, KS = S::KS;
  constexpr int NF = nfields_of(EQ);
  constexpr int WAVES = mfma_waves_of(EQ, ORDER);
  constexpr int BLOCK = 64 * WAVES;

  constexpr bool SH = short_last_block(NP);   // last output block on 4x4x4
  constexpr bool DF = direct_flux<EQ, ORDER>();
  constexpr bool RN = ref_normal_lift<EQ, ORDER>();
  constexpr int KSD = direct_ksteps_of(EQ, ORDER);   // direct flux: LIFT k-steps
  constexpr int LT = DF ? MB * KSD : S::LIFT_TILES;
  constexpr int C = flux_chunk<EQ, ORDER>();
  constexpr int NCHUNK = (NFP + C - 1) / C;
  constexpr int FROW = kFluxRow;
  constexpr bool ST = stage_own_traces<EQ, ORDER>();
  constexpr bool FG = face_geometry_in_lds<EQ, ORDER>() && !(POST && DF && ST);
  // LDS bytes of the kernel with element row pitch sp
  constexpr auto lds_bytes = [](int sp) {
    return 8 * (S::OP_TILES * 64 + LT * 64 + (POST ? MB * KS * 64 : 1)) + 4 * NFN +
           8 * WAVES * (wave_buffer<EQ, ORDER>(sp) + (ST ? NF * 16 * sp : 1)) +
           16 * WAVES * (FG ? 128 : 1);
  };
  // the odd pitch where it leaves the blocks per CU that the LDS allows
  constexpr int SP = kLdsPerCu / lds_bytes(stage_pitch(NP)) == kLdsPerCu / lds_bytes(NP)
                         ? stage_pitch(NP) : NP;
  constexpr int WB = wave_buffer<EQ, ORDER>(SP);

  __shared__ double s_op[S::OP_TILES * 64];
  __shared__ double s_lift[LT * 64];
  __shared__ double s_post[POST ? MB * KS * 64 : 1];
  __shared__ int s_fmask[NFN];
  // per wave: scaled flux of C LIFT k-steps in MFMA B order [kk][field][FROW],
  // then one field of the output tile
  __shared__ double s_buf[WAVES][WB];
  // per wave, if staged: the tile's state [field][n * SP + j]
  __shared__ double s_stage[WAVES][ST ? NF * 16 * SP : 1];
  double* const stg = s_stage[threadIdx.x >> 6];
  // per wave, if enabled: the tile's face geometry [face * 16 + n][4] (not
  // with the post-multiplication on top of the staged state, where its 8 KB
  // would cost a block per CU at wave P3)
  __shared__ double2 s_fgeo[WAVES][FG ? 128 : 1];
  double2* const fgl = s_fgeo[threadIdx.x >> 6];

  const int tid = threadIdx.x;
  for (int k = tid; k < S::OP_TILES * 64; k += BLOCK) s_op[k] = a.op_tiles[k];
  for (int k = tid; k < LT * 64; k += BLOCK) s_lift[k] = a.lift_tiles[k];
  for (int k = tid; k < NFN; k += BLOCK) s_fmask[k] = a.fmask[k];
  if constexpr (POST)
    for (int k = tid; k < MB * KS * 64; k += BLOCK) s_post[k] = a.post_tiles[k];
  __syncthreads();

  const int lane = tid & 63;
  const int g = lane >> 4;
  const int n = lane & 15;
  // every wave is a worker of its XCD's slice of the tile schedule; waves of
  // a block are independent
  const int slot = block_in_xcd() * WAVES + (tid >> 6);
  const int nslots = blocks_per_xcd() * WAVES;
  const int32_t p_end = a.sched.part[block_xcd() + 1];
  const double c0 = a.p0;
  double* const buf = (DF && ST) ? stg : s_buf[tid >> 6];
  // in-tile neighbours: buffer resources of the state fields (each below 4
  // GiB, which the plan checks)
  constexpr bool IT = intile_neighbours_of(EQ, ORDER);
  const uint32_t field_bytes = (uint32_t)(a.nelem * NP * 8);
  __amdgpu_buffer_rsrc_t rsrc[IT ? NF : 1];
  if constexpr (IT) {
#pragma unroll
    for (int c = 0; c < NF; ++c)
      rsrc[c] = __builtin_amdgcn_make_buffer_rsrc((void*)a.q.p[c], (short)0, field_bytes,
                                                  kBufferWord3);
  }

  // exterior-map entries of one flux chunk's items, loaded one chunk ahead
  // (the first chunk's before the volume phase) so that the gathers do not
  // wait on the map
  constexpr bool PF = !DF && prefetch_ext<EQ, ORDER>();
  int32_t ext_pf[C];
  auto load_ext = [&](int64_t t, int ch) {
#pragma unroll
    for (int it = 0; it < C; ++it)
      ext_pf[it] = a.ext_t[t * flux_items_of(EQ, ORDER) + (ch * C + it) * 64 + lane];
  };

  // the next tile's index is loaded while this tile runs (sched_at: through
  // the scalar cache)
  int32_t pos = __builtin_amdgcn_readfirstlane(a.sched.part[block_xcd()] + slot);
  int32_t next_tile = pos < p_end ? sched_at(a.sched.order, pos) : 0;

  // volume operands of a tile: state at the lane's B-operand nodes
  // j = g * KS + ks of element n (0 past Np), and rx
  constexpr bool PV = pipe_volume<EQ, ORDER>();
  double qv[PV ? NF : 1][PV ? KS : 1];
  double rxv[PV ? 9 : 1];
  auto load_volume = [&](int64_t t) {
    const int64_t ge = t * 16 + n;
    const int64_t e = ge < a.nelem ? ge : a.nelem - 1;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int c = 0; c < NF; ++c)
        qv[c][ks] = g * KS + ks < NP ? a.q.p[c][e * NP + g * KS + ks] : 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) rxv[k] = a.vgeo[(t * 9 + k) * 16 + n];
  };
  if constexpr (PV)
    if (pos < p_end) load_volume(next_tile);

  for (; pos < p_end; pos += nslots) {
    const int64_t tile = next_tile;
    if (pos + nslots < p_end) next_tile = sched_at(a.sched.order, pos + nslots);
    // direct flux: every k-step's exterior-map entry, loaded with the volume
    // operands so that no gather waits on the map (wave: -1.5%; it costs
    // registers that Maxwell P4 cannot spare, +9%)
    constexpr bool MAP0 = DF && EQ == DG_EQ_WAVE;
    int32_t ext_all[MAP0 ? KSD : 1];
    if constexpr (MAP0) {
#pragma unroll
      for (int ks = 0; ks < KSD; ++ks) ext_all[ks] = a.ext_d[(tile * KSD + ks) * 64 + lane];
    }
    // the tile's face geometry into LDS: formed from rx where the plan found
    // Fscale n = rx^T c_f on every face (always with the reference-normal
    // LIFT, which needs it), else copied from the plan's table
    constexpr bool FGRX = FG && !(EQ == DG_EQ_MAXWELL && ORDER == 4);
    const bool geo_rx = FGRX && (RN || a.face_geo_rx != 0);
    if constexpr (FG) {
      if (!geo_rx) {
        const double2* src = reinterpret_cast<const double2*>(a.fgeo) + tile * 128;
        const double2 f0 = src[lane], f1 = src[lane + 64];
        fgl[lane] = f0;
        fgl[lane + 64] = f1;
      }
    }
    if constexpr (PF) load_ext(tile, 0);
    const int64_t gel = tile * 16 + n;
    const bool valid = gel < a.nelem;
    const int64_t el = valid ? gel : a.nelem - 1;   // clamped for loads
    double rx[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) rx[k] = PV ? rxv[k] : a.vgeo[(tile * 9 + k) * 16 + n];
    if constexpr (FGRX) if (geo_rx) {
      // The tile's face geometry from rx instead of the plan's table: lane
      // (g, n) forms face g of element n as Fscale n = rx^T c_g (within
      // rounding of the table's values).  Measured on MI355X: wave P3 -3%
      // (162k tets), -3.9% (216k tets), wave P4 -0.8%, Euler P1 -10%, P2
      // -5.5%, P3-P4 -1%, Maxwell P2 -12.5%, P3 -10%; Maxwell P4 +1.2%, so
      // it copies the table there.
      double m[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        const double c0 = g == 3 ? -1.0 : (g == 2 ? 1.0 : 0.0);
        const double c1 = g == 1 ? -1.0 : (g == 2 ? 1.0 : 0.0);
        const double c2 = g == 0 ? -1.0 : (g == 2 ? 1.0 : 0.0);
        m[x] = c0 * rx[3 * x] + c1 * rx[3 * x + 1] + c2 * rx[3 * x + 2];
      }
      const double fs = sqrt_nr(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      const double inv = rcp_nr(fs);
      fgl[2 * (g * 16 + n)] = double2{m[0] * inv, m[1] * inv};
      fgl[2 * (g * 16 + n) + 1] = double2{m[2] * inv, fs};
    }
    // state at volume B-operand node j = g * KS + ks (callers check j < NP)
    auto qload = [&](int c, int ks) -> double {
      if constexpr (PV) return qv[c][ks];
      else return a.q.p[c][el * NP + g * KS + ks];
    };

    f64x4 acc[NF][MB];
#pragma unroll
    for (int c = 0; c < NF; ++c)
#pragma unroll
      for (int b = 0; b < MB; ++b) acc[c][b] = f64x4{0.0, 0.0, 0.0, 0.0};

    // {{{ volume terms

    // wave: the gradient's per-r accumulators t_r = D_r u (with the
    // reference-normal LIFT they take the velocity flux too, and are combined
    // with rx after the surface phase)
    f64x4 t[EQ == DG_EQ_WAVE ? 3 : 1][MB];
    if constexpr (EQ == DG_EQ_WAVE) {
      // u_t = -c sum_r D_r w_r (w_r = sum_x rx[x][r] v_x); v_x,t = -c sum_r rx[x][r] D_r u;
      // the A tiles hold -c D_r
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int b = 0; b < MB; ++b) t[r][b] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll KUNROLL
      for (int ks = 0; ks < KS; ++ks) {
        const int j = g * KS + ks;
        double u = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
        if (j < NP) {
          u = qload(0, ks);
          vx = qload(1, ks);
          vy = qload(2, ks);
          vz = qload(3, ks);
          if constexpr (ST) {
            stg[0 * 16 * SP + n * SP + j] = u;
            stg[1 * 16 * SP + n * SP + j] = vx;
            stg[2 * 16 * SP + n * SP + j] = vy;
            stg[3 * 16 * SP + n * SP + j] = vz;
          }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const double w = rx[r] * vx + rx[3 + r] * vy + rx[6 + r] * vz;
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const double av = s_op[((r * MB + b) * KS + ks) * 64 + lane];
            mma(t[r][b], av, u, SH && b == MB - 1);
            mma(acc[0][b], av, w, SH && b == MB - 1);
          }
        }
      }
      // (the host's D tiles of the wave operator carry the factor -c)
      if constexpr (!RN) {
#pragma unroll
        for (int b = 0; b < MB; ++b)
#pragma unroll
          for (int v = 0; v < ncomp(SH && b == MB - 1); ++v)
#pragma unroll
            for (int x = 0; x < 3; ++x)
              acc[1 + x][b][v] = rx[3 * x] * t[0][b][v] + rx[3 * x + 1] * t[1][b][v] +
                                 rx[3 * x + 2] * t[2][b][v];
      }
    } else if constexpr (EQ == DG_EQ_EULER) {
      // q_t = sum_r D_r w_{c,r},  w_{c,r} = sum_x rx[x][r] F_{c,x}(q)
      const double gamma = c0;
#pragma unroll KUNROLL
      for (int ks = 0; ks < KS; ++ks) {
        const int j = g * KS + ks;
        double w[5][3];
        if (j < NP) {
          const double rho = qload(0, ks), e = qload(1, ks);
          const double m[3] = {qload(2, ks), qload(3, ks), qload(4, ks)};
          if constexpr (ST) {
            stg[0 * 16 * SP + n * SP + j] = rho;
            stg[1 * 16 * SP + n * SP + j] = e;
#pragma unroll
            for (int x = 0; x < 3; ++x) stg[(2 + x) * 16 * SP + n * SP + j] = m[x];
          }
          const double inv_rho = rcp_nr(rho);
          const double u[3] = {m[0] * inv_rho, m[1] * inv_rho, m[2] * inv_rho};
          const double p = (gamma - 1.0) * (e - 0.5 * (m[0] * u[0] + m[1] * u[1] + m[2] * u[2]));
          const double h = (e + p) * inv_rho;
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
              w[c][r] = rx[r] * f[c][0] + rx[3 + r] * f[c][1] + rx[6 + r] * f[c][2];
        } else {
#pragma unroll
          for (int c = 0; c < 5; ++c) w[c][0] = w[c][1] = w[c][2] = 0.0;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const double av = s_op[((r * MB + b) * KS + ks) * 64 + lane];
#pragma unroll
            for (int c = 0; c < 5; ++c) mma(acc[c][b], av, w[c][r], SH && b == MB - 1);
          }
      }
    } else {
      // E_t = curl H, H_t = -curl E in strong form.  rx is constant per
      // element, so the Levi-Civita combination moves into the B operands:
      //   (curl F)_o = sum_{x,a} eps_oxa d_x F_a = sum_r D_r w_{o,r},
      //   w_{o,r} = sum_{x,a} eps_oxa rx[x][r] F_a,
      // and every output field is one contraction per r, as in the
      // divergence form (no per-field derivative accumulators).
#pragma unroll KUNROLL
      for (int ks = 0; ks < KS; ++ks) {
        const int j = g * KS + ks;
        double f[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) f[c] = 0.0;
        if (j < NP) {
#pragma unroll
          for (int c = 0; c < 6; ++c) f[c] = qload(c, ks);
          if constexpr (ST)
#pragma unroll
            for (int c = 0; c < 6; ++c) stg[c * 16 * SP + n * SP + j] = f[c];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          double w[6];
#pragma unroll
          for (int o = 0; o < 3; ++o) {
            double ch = 0.0, ce = 0.0;
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const int s = levi(o, x, c);
                if (s == 0) continue;
                ch += (double)s * rx[3 * x + r] * f[3 + c];
                ce += (double)s * rx[3 * x + r] * f[c];
              }
            w[o] = ch;         // E rhs: + curl H
            w[3 + o] = -ce;    // H rhs: - curl E
          }
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const double av = s_op[((r * MB + b) * KS + ks) * 64 + lane];
#pragma unroll
            for (int c = 0; c < 6; ++c) mma(acc[c][b], av, w[c], SH && b == MB - 1);
          }
        }
      }
    }

    // }}}

    // {{{ surface terms

    if constexpr (ST || FG) wave_lds_fence();   // staged state / face geometry complete

    if constexpr (DF) {
      // direct: at k-step ks, lane (g, n) evaluates the flux at face-node pair
      // direct_pair_of(ks, g) of element n and feeds it to the LIFT MFMAs.
      // One k-step's flux: pair fp (padded slots read the face's node 0, a
      // valid node, and contribute nothing), its face f, and the state traces.
      auto trace_pair = [&](int ks, int& f, bool& fok, int64_t& vn, double2& g01, double2& g23,
                            double* qm, double* qp) {
        const int fp = direct_pair_of(EQ, ORDER, ks, g);
        fok = fp >= 0;
        f = fok ? fp / NFP : ks / ((NFP + 3) / 4);
        const int vloc = s_fmask[fok ? fp : f * NFP];
        vn = el * NP + vloc;
        const double2* pgeo = FG ? fgl + 2 * (f * 16 + n)
                                 : reinterpret_cast<const double2*>(a.fgeo + (tile * 64 + f * 16 + n) * 4);
        g01 = pgeo[0];
        g23 = pgeo[1];
        if constexpr (ST) {
#pragma unroll
          for (int c = 0; c < NF; ++c) qm[c] = stg[c * 16 * SP + n * SP + vloc];
        } else {
#pragma unroll
          for (int c = 0; c < NF; ++c) qm[c] = a.q.p[c][vn];
        }
        int32_t ext;
        if constexpr (MAP0) ext = ext_all[ks];
        else ext = a.ext_d[(tile * KSD + ks) * 64 + lane];
        if constexpr (IT) {
          // in-tile neighbours from the staged state, the others by buffer
          // loads (an offset past the end for the lanes that need none); a
          // boundary node takes its own trace
          const bool loc = ext >= kInTile;
          const bool glob = ext >= 0 && !loc;
          // the staged trace: the in-tile neighbour's, else the lane's own
          // (the boundary state's base)
          constexpr bool BLEND = !(EQ == DG_EQ_WAVE && ORDER == 2);
          const int lofs = loc ? ((ext >> 8) & 15) * SP + (ext & 255) : (BLEND ? n * SP + vloc : 0);
          const uint32_t boff = glob ? (uint32_t)ext * 8u : field_bytes;
          // qp = gv + wl lv: the buffer load gives exactly 0 where it is out of
          // range, and both loads stay unconditional.  A select lets the
          // compiler move each load into a branch of its own, one round trip
          // each: measured on MI355X, the blend is 9% faster at wave P3 (5%
          // on 384k tets) and the select 5-9% faster at wave P2, where the
          // blend's registers cost a wave per SIMD.
          const double wl = glob ? 0.0 : 1.0;
#pragma unroll
          for (int c = 0; c < NF; ++c) {
            const double gv = __builtin_bit_cast(
                double, __builtin_amdgcn_raw_buffer_load_b64(rsrc[c], boff, 0, 0));
            const double lv = stg[c * 16 * SP + lofs];
            qp[c] = BLEND ? __builtin_fma(lv, wl, gv) : (glob ? gv : (loc ? lv : qm[c]));
          }
        } else {
          // a boundary node loads its own trace (src = vn), which is the
          // boundary state up to the flux's per-field signs where it has them
          const int64_t src = ext >= 0 ? (int64_t)ext : vn;
#pragma unroll
          for (int c = 0; c < NF; ++c) qp[c] = a.q.p[c][src];
        }
        if constexpr (lean_flux<EQ, ORDER>()) {
#pragma unroll
          for (int c = 0; c < NF; ++c)
            if (Flux::mirror_sign(c) < 0) qp[c] = ext >= 0 ? qp[c] : -qp[c];
        } else {
          double qb[NF];
          const double nrm[3] = {g01.x, g01.y, g23.x};
          Flux::boundary(qm, nrm, qb);
#pragma unroll
          for (int c = 0; c < NF; ++c) qp[c] = ext >= 0 ? qp[c] : qb[c];
        }
      };
      if constexpr (RN) {
        // wave, velocity flux LIFTed in reference normals into t_r (every
        // k-step unrolled, so the faces -- and the t_r -- a k-step touches
        // are known at compile time)
#pragma unroll
        for (int ks = 0; ks < KSD; ++ks) {
          int f;
          bool fok;
          int64_t vn;
          double2 g01, g23;
          double qm[4], qp[4];
          trace_pair(ks, f, fok, vn, g01, g23, qm, qp);
          // the flux's u part times Fscale, and s of its velocity part s n
          // (c_{f,r} of this lane's face applied per t_r below)
          double bu, bs;
          if constexpr (lean_flux<EQ, ORDER>()) {
            // with a = v-.n, b = v+.n and h = c / 2 (wave_flux): u part
            // h (a + b + u+ - u-), s = h (u- + u+ + b - a)
            const double am = qm[1] * g01.x + qm[2] * g01.y + qm[3] * g23.x;
            const double bp = qp[1] * g01.x + qp[2] * g01.y + qp[3] * g23.x;
            const double h = 0.5 * c0;
            bu = (h * g23.y) * ((am + bp) + (qp[0] - qm[0]));
            bs = h * ((qm[0] + qp[0]) + (bp - am));
          } else {
            // wave_flux as written: c ({v}.n + [u]/2) and s = c ({u} + [v].n/2)
            const double uavg = 0.5 * (qm[0] + qp[0]);
            const double udiff = qp[0] - qm[0];
            const double vavg_n = 0.5 * ((qm[1] + qp[1]) * g01.x + (qm[2] + qp[2]) * g01.y +
                                         (qm[3] + qp[3]) * g23.x);
            const double vjump = (qp[1] - qm[1]) * g01.x + (qp[2] - qm[2]) * g01.y +
                                 (qp[3] - qm[3]) * g23.x;
            bu = g23.y * (c0 * (vavg_n + 0.5 * udiff));
            bs = c0 * (uavg + 0.5 * vjump);
          }
          bu = fok ? bu : 0.0;
          bs = fok ? bs : 0.0;
          // faces this k-step covers, over its four lane groups
          constexpr int FPK = (NFP + 3) / 4;
          const int f_lo = direct_pair_of(EQ, ORDER, ks, 0) >= 0
                               ? direct_pair_of(EQ, ORDER, ks, 0) / NFP : ks / FPK;
          int f_hi = f_lo;
#pragma unroll
          for (int gg = 1; gg < 4; ++gg)
            if (direct_pair_of(EQ, ORDER, ks, gg) >= 0)
              f_hi = direct_pair_of(EQ, ORDER, ks, gg) / NFP;
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const bool sh = SH && b == MB - 1;
            const double av = s_lift[(b * KSD + ks) * 64 + lane];
            mma(acc[0][b], av, bu, sh);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              bool any = false;
#pragma unroll
              for (int ff = 0; ff < 4; ++ff) any |= ff >= f_lo && ff <= f_hi && ref_face_coef(ff, r) != 0;
              if (any) {
                const double cs = (double)ref_face_coef(f, r) * bs;
                mma(t[r][b], av, cs, sh);
              }
            }
          }
        }
#pragma unroll
        for (int b = 0; b < MB; ++b)
#pragma unroll
          for (int v = 0; v < ncomp(SH && b == MB - 1); ++v)
#pragma unroll
            for (int x = 0; x < 3; ++x)
              acc[1 + x][b][v] = rx[3 * x] * t[0][b][v] + rx[3 * x + 1] * t[1][b][v] +
                                 rx[3 * x + 2] * t[2][b][v];
      } else {
#pragma unroll(ORDER == 4 ? DUNROLL_P4 : DUNROLL)
        for (int ks = 0; ks < KSD; ++ks) {
          int f;
          bool fok;
          int64_t vn;
          double2 g01, g23;
          double qm[NF], qp[NF], out[NF];
          trace_pair(ks, f, fok, vn, g01, g23, qm, qp);
          const double nrm[3] = {g01.x, g01.y, g23.x};
          Flux::flux(qm, qp, nrm, c0, out);
          double bv[NF];
#pragma unroll
          for (int c = 0; c < NF; ++c) bv[c] = fok ? g23.y * out[c] : 0.0;
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const double av = s_lift[(b * KSD + ks) * 64 + lane];
#pragma unroll
            for (int c = 0; c < NF; ++c) mma(acc[c][b], av, bv[c], SH && b == MB - 1);
          }
        }
      }
      // the output tile goes through the stage (buf == stg): every lane's
      // reads of the state are done
      if constexpr (ST) wave_lds_fence();
    } else {
    // flux chunks unrolled by 2 at Maxwell P3 (measured -6%; Euler P3 +12%,
    // and +10% in the fused R M form of the adjoint, +8% with the fused stage)
#pragma unroll((EQ == DG_EQ_MAXWELL && ORDER == 3 && !POST && !STAGE) ? 2 : 1)
    for (int ch = 0; ch < NCHUNK; ++ch) {
      // producer: items (face pair p, node kk) with kk fastest
#pragma unroll
      for (int it = 0; it < C; ++it) {
        const int item = it * 64 + lane;
        const int kk = item % C;
        const int p = item / C;                   // p = face * 16 + element
        const int ks = ch * C + kk;
        if (ks < NFP) {
          const int fg = p >> 4, fe = p & 15;
          int64_t pe = tile * 16 + fe;
          pe = pe < a.nelem ? pe : a.nelem - 1;
          const double2* pgeo = FG ? fgl + 2 * p
                                   : reinterpret_cast<const double2*>(a.fgeo + (tile * 64 + p) * 4);
          const double2 g01 = pgeo[0], g23 = pgeo[1];
          const double nrm[3] = {g01.x, g01.y, g23.x};
          const double fscale = g23.y;
          const int fn = fg * NFP + ks;
          const int64_t vn = pe * NP + s_fmask[fn];
          double qm[NF], qp[NF], out[NF];
          if constexpr (ST) {
#pragma unroll
            for (int c = 0; c < NF; ++c) qm[c] = stg[c * 16 * SP + fe * SP + s_fmask[fn]];
          } else {
#pragma unroll
            for (int c = 0; c < NF; ++c) qm[c] = a.q.p[c][vn];
          }
          const int32_t ext =
              PF ? ext_pf[it] : a.ext_t[tile * flux_items_of(EQ, ORDER) + (ch * C + it) * 64 + lane];
          if constexpr (branch_free_gathers<EQ, ORDER>()) {
            const int64_t src = ext >= 0 ? (int64_t)ext : vn;
#pragma unroll
            for (int c = 0; c < NF; ++c) qp[c] = a.q.p[c][src];
            if constexpr (lean_flux<EQ, ORDER>()) {
#pragma unroll
              for (int c = 0; c < NF; ++c)
                if (Flux::mirror_sign(c) < 0) qp[c] = ext >= 0 ? qp[c] : -qp[c];
            } else {
              double qb[NF];
              Flux::boundary(qm, nrm, qb);
#pragma unroll
              for (int c = 0; c < NF; ++c) qp[c] = ext >= 0 ? qp[c] : qb[c];
            }
          } else if (ext >= 0) {
#pragma unroll
            for (int c = 0; c < NF; ++c) qp[c] = a.q.p[c][ext];
          } else {
            Flux::boundary(qm, nrm, qp);
          }
          Flux::flux(qm, qp, nrm, c0, out);
#pragma unroll
          for (int c = 0; c < NF; ++c) buf[(kk * NF + c) * FROW + p] = fscale * out[c];
        }
      }
      if constexpr (PF)
        if (ch + 1 < NCHUNK) load_ext(tile, ch + 1);
      wave_lds_fence();
      // consumer
#pragma unroll
      for (int kk = 0; kk < C; ++kk) {
        const int ks = ch * C + kk;
        if (ks < NFP) {
          double bv[NF];
#pragma unroll
          for (int c = 0; c < NF; ++c) bv[c] = buf[(kk * NF + c) * FROW + lane];
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const double av = s_lift[(b * NFP + ks) * 64 + lane];
#pragma unroll
            for (int c = 0; c < NF; ++c) mma(acc[c][b], av, bv[c], SH && b == MB - 1);
          }
        }
      }
      wave_lds_fence();
    }
    }

    // }}}

    // {{{ output, one field at a time through the buffer to contiguous stores
    // (non-temporal: the RHS is not read again by this kernel, and keeping it
    // out of L2 leaves room for the state the neighbours' gathers reuse;
    // unrolled, so that a field's LDS reads are in flight together: measured
    // Maxwell P3 -9%, wave P4 -6%, Euler P2 -4%)

    // the next tile's volume operands, ahead of this tile's stores
    if constexpr (PV)
      if (pos + nslots < p_end) load_volume(next_tile);

    // STAGE: each lane's slots of the output tile, idx = lane + 64 t
    constexpr int NSLOT = (16 * NP + 63) / 64;
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      // STAGE: this field's res and q_in loads are issued before the tile
      // goes through LDS, so they are in flight together
      double st_res[STAGE ? NSLOT : 1], st_q[STAGE ? NSLOT : 1];
      if constexpr (STAGE) {
        const int64_t off = tile * 16 * NP;
#pragma unroll
        for (int t = 0; t < NSLOT; ++t) {
          const int idx = lane + 64 * t;
          const bool ok = idx < 16 * NP && tile * 16 + idx / NP < a.nelem;
          // a == 0 (the first stage) does not read res, so it may start uninitialised
          st_res[t] = ok && a.stage_a != 0.0 ? a.stage_res.p[c][off + idx] : 0.0;
          st_q[t] = ok ? a.q.p[c][off + idx] : 0.0;
        }
      }
#pragma unroll
      for (int b = 0; b < MB; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * b + g + 4 * v;
          if (i < NP) buf[n * SP + i] = acc[c][b][v];
        }
      wave_lds_fence();
      if constexpr (POST) {
        // out[i][n] = sign_c |J_n| sum_j M_ref[i][j] out[j][n]: B operand
        // node j = g * KS + ks of element n, read back from the buffer
        f64x4 pacc[MB];
#pragma unroll
        for (int b = 0; b < MB; ++b) pacc[b] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int j = g * KS + ks;
          const double bv = j < NP ? buf[n * SP + j] : 0.0;
#pragma unroll
          for (int b = 0; b < MB; ++b)
            mma(pacc[b], s_post[(b * KS + ks) * 64 + lane], bv, SH && b == MB - 1);
        }
        wave_lds_fence();
        const double sc = a.post_sign[c] * a.jac[el];
#pragma unroll
        for (int b = 0; b < MB; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = 16 * b + g + 4 * v;
            if (i < NP) buf[n * SP + i] = sc * pacc[b][v];
          }
        wave_lds_fence();
      }
      if constexpr (STAGE) {
        const int64_t off = tile * 16 * NP;
        double* res = a.stage_res.p[c] + off;
        double* qo = a.stage_qout.p[c] + off;
#pragma unroll
        for (int t = 0; t < NSLOT; ++t) {
          const int idx = lane + 64 * t;
          if (idx < 16 * NP && tile * 16 + idx / NP < a.nelem) {
            const double r = (a.stage_a != 0.0 ? a.stage_a * st_res[t] : 0.0) + a.stage_dt * buf[(idx / NP) * SP + idx % NP];
            res[idx] = r;
            qo[idx] = a.stage_beta * st_q[t] + a.stage_b * r;
          }
        }
      } else {
        // all of the lane's slots read from LDS, then stored (non-temporal
        // stores: ordinary ones measured 5-50% slower)
        double* dst = a.rhs.p[c] + tile * 16 * NP;
        double vals[NSLOT];
#pragma unroll
        for (int t = 0; t < NSLOT; ++t) {
          const int idx = lane + 64 * t;
          vals[t] = idx < 16 * NP ? buf[(idx / NP) * SP + idx % NP] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < NSLOT; ++t) {
          const int idx = lane + 64 * t;
          if (idx < 16 * NP && tile * 16 + idx / NP < a.nelem)
            __builtin_nontemporal_store(vals[t], dst + idx);
        }
      }
      wave_lds_fence();
    }

    // }}}
  }
}

}  // namespace dg
