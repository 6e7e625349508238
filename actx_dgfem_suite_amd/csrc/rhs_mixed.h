// Mixed-precision fused RHS kernel: FP32 state in and RHS out, the D and LIFT
// contractions on the matrix cores in FP32 (v_mfma_f32_16x16x4f32) or as a
// three-product bf16 split (gfx950's v_mfma_f32_16x16x32_bf16, or
// v_mfma_f32_16x16x16bf16_1k where the wider form's padding costs occupancy)
// with FP32 accumulation, and everything else -- geometric factors, the numerical flux
// (Euler pressure recovery included) and the combination of reference
// derivatives into physical ones -- in FP64 on values promoted from FP32.
//
// The work decomposition is the FP64 kernel's (rhs_mfma.h): each wave owns
// tiles of 16 elements (the MFMA N dimension), B operands come straight from
// HBM, the flux is produced in a gather-friendly order and handed over
// through a small per-wave LDS buffer, and outputs leave through the same
// buffer as contiguous stores.
//
// Operand layouts (lane l, g = l / 16, n = l % 16), measured on gfx950:
//   f32 16x16x4:    A[m = n][k = g], B[k = g][col = n], C[row = 4 g + v][col = n]
//   bf16 16x16x32:  lane l supplies 8 consecutive k = 8 g + i (i = 0..7) of
//                   A[m = n][k] and B[k][col = n]; C as for f32.
// Contraction index k at FP32 k-step ks stands for node j = g * KS + ks (a
// contiguous run per lane group, as in the FP64 kernel); the bf16 form takes
// eight consecutive k-steps per instruction (KS rounded up to a multiple of
// 8).  On MI355X the 16x16x32 form runs 1.94 PFLOP/s dense against 1.06 for
// CDNA3's 16x16x16_1k (same time per instruction, twice the k).
//
// BF16x3: x = x_hi + x_lo with x_hi = bf16(x), x_lo = bf16(x - x_hi); the
// product is a_hi b_hi + a_hi b_lo + a_lo b_hi (the a_lo b_lo term, ~2^-16
// relative, is dropped).  A operands are split on the host, B on the fly.
#pragma once

#include "dg_common.h"
#include "rhs_mfma.h"

namespace dg {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

struct MixedArgs {
  Schedule sched;           // 16-element tiles
  const float* q[kMaxFields];
  float* rhs[kMaxFields];
  const double* geo;        // (K, 25), FP64
  const int32_t* emap;      // (K, 4*Nfp)
  const int32_t* fmask;     // (4*Nfp)
  // the FP64 MFMA kernel's tile-ordered tables (rhs_mfma.h): rx, face
  // geometry, and the exterior map in that kernel's flux-item order
  const double* vgeo;       // [tile][9][16]
  const double* fgeo;       // [tile][face*16 + n][4]
  int32_t face_geo_rx;      // 1: every face satisfies Fscale n = rx^T c_f
  const int32_t* ext_t;     // [tile][chunk][item][64]
  const int32_t* ext_d;     // [tile][kstep][64]: the direct flux order (in-tile entries marked)
  const void* op_tiles;     // A tiles of D (float, or bf16 hi/lo pairs)
  const void* lift_tiles;   // A tiles of LIFT
  int64_t nelem;
  double p0;
  // STAGE only: the LSERK4 stage applied to the RHS tile -- the FP64 state
  // and residual of the tile's own elements are updated in place, and the
  // new FP32 shadow goes to st_q32 (not q, whose neighbour traces other
  // waves may still gather)
  double* st_q[kMaxFields];
  double* st_res[kMaxFields];
  float* st_q32[kMaxFields];
  double st_a, st_b, st_dt;
};

// round-to-nearest-even float -> bf16 bits (finite inputs)
__host__ __device__ inline uint16_t bf16_bits(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

// bf16 k-steps per instruction of the BF16x3 form: 8 on gfx950's
// v_mfma_f32_16x16x32_bf16 or 4 on v_mfma_f32_16x16x16bf16_1k.  The 16x16x32
// form runs twice the flops per instruction (1.94 against 1.06 PFLOP/s
// measured on MI355X), but pads k to 8 per lane group in the A tiles, the
// flux buffer and the on-the-fly B split, and the contractions are not what
// bounds these kernels: measured, it is 1.5% faster at Maxwell P4 and even to
// 45% slower elsewhere (wave P2 +22%, Euler P1 +43%, Maxwell P1 +8%).
__host__ __device__ constexpr int bf16_kgroup_of(int eq, int order) {
  return (eq == DG_EQ_MAXWELL && order == 4) ? 8 : 4;
}

// Whether the B split uses gfx950's v_cvt_pk_bf16_f32 (two values per
// instruction) rather than integer rounding (bf16_bits, about 11 VALU
// operations per value).  Measured on MI355X: Euler P1 -8%, P2 -5%, P4 -5%,
// Maxwell P1 -6%, wave even; at Maxwell P2 it costs a wave per SIMD (+37%).
__host__ __device__ constexpr bool bf16_hw_split_of(int eq, int order) {
  return !(eq == DG_EQ_MAXWELL && order == 2);
}

template <int PREC, int KGB, bool HW = true>
struct MixedMma;

// FP32: one k per lane per instruction
template <int KGB, bool HW>
struct MixedMma<DG_PREC_FP32, KGB, HW> {
  static constexpr int KG = 1;                  // k-steps per instruction
  static constexpr int A_WORDS = 1;             // 32-bit words of A per lane
  __device__ static int row(int g, int v) { return 4 * g + v; }
  __device__ static f32x4 mma(const uint32_t* a, const float* b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a[0]), b[0], c, 0, 0, 0);
  }
};

// BF16x3 split of the B operand, KG values -> packed hi and lo words: the
// conversions are gfx950's v_cvt_pk_bf16_f32 (round to nearest even, as the
// host's split of A), two values per instruction
template <int KG>
__device__ inline void split_bf16(const float* b, uint32_t* hw, uint32_t* lw) {
#pragma unroll
  for (int i = 0; i < KG / 2; ++i) {
    const bf16x2 h = {(__bf16)b[2 * i], (__bf16)b[2 * i + 1]};
    const bf16x2 l = {(__bf16)(b[2 * i] - (float)h[0]), (__bf16)(b[2 * i + 1] - (float)h[1])};
    hw[i] = __builtin_bit_cast(uint32_t, h);
    lw[i] = __builtin_bit_cast(uint32_t, l);
  }
}

// BF16x3 on v_mfma_f32_16x16x32_bf16: eight k per lane per instruction, A as
// (hi[8], lo[8]) bf16 in 8 words
template <bool HW>
struct MixedMma<DG_PREC_BF16X3, 8, HW> {
  static constexpr int KG = 8;
  static constexpr int A_WORDS = 8;             // hi: 4 words, lo: 4 words
  __device__ static int row(int g, int v) { return 4 * g + v; }
  __device__ static f32x4 mma(const uint32_t* a, const float* b, f32x4 c) {
    uint32_t hw[4], lw[4];
    static_assert(HW, "the 16x16x32 form splits with v_cvt_pk_bf16_f32");
    split_bf16<8>(b, hw, lw);
    const bf16x8 ahi = __builtin_bit_cast(bf16x8, (u32x4){a[0], a[1], a[2], a[3]});
    const bf16x8 alo = __builtin_bit_cast(bf16x8, (u32x4){a[4], a[5], a[6], a[7]});
    const bf16x8 bhi = __builtin_bit_cast(bf16x8, (u32x4){hw[0], hw[1], hw[2], hw[3]});
    const bf16x8 blo = __builtin_bit_cast(bf16x8, (u32x4){lw[0], lw[1], lw[2], lw[3]});
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi, c, 0, 0, 0);
  }
};

// BF16x3 on v_mfma_f32_16x16x16bf16_1k: four k per lane per instruction, A as
// (hi[4], lo[4]) bf16 in 4 words; B split by v_cvt_pk_bf16_f32
template <>
struct MixedMma<DG_PREC_BF16X3, 4, true> {
  static constexpr int KG = 4;
  static constexpr int A_WORDS = 4;             // hi: 2 words, lo: 2 words
  __device__ static int row(int g, int v) { return 4 * g + v; }
  __device__ static f32x4 mma(const uint32_t* a, const float* b, f32x4 c) {
    uint32_t hw[2], lw[2];
    split_bf16<4>(b, hw, lw);
    const bf16x4 ahi = __builtin_bit_cast(bf16x4, (u32x2){a[0], a[1]});
    const bf16x4 alo = __builtin_bit_cast(bf16x4, (u32x2){a[2], a[3]});
    const bf16x4 bhi = __builtin_bit_cast(bf16x4, (u32x2){hw[0], hw[1]});
    const bf16x4 blo = __builtin_bit_cast(bf16x4, (u32x2){lw[0], lw[1]});
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, c, 0, 0, 0);
  }
};

// The same with the B split by integer rounding, operand by operand (the
// form that keeps Maxwell P2 at two waves per SIMD)
template <>
struct MixedMma<DG_PREC_BF16X3, 4, false> {
  static constexpr int KG = 4;
  static constexpr int A_WORDS = 4;
  __device__ static int row(int g, int v) { return 4 * g + v; }
  __device__ static f32x4 mma(const uint32_t* a, const float* b, f32x4 c) {
    bf16x4 ahi, alo, bhi, blo;
    ahi[0] = (short)(a[0] & 0xffff); ahi[1] = (short)(a[0] >> 16);
    ahi[2] = (short)(a[1] & 0xffff); ahi[3] = (short)(a[1] >> 16);
    alo[0] = (short)(a[2] & 0xffff); alo[1] = (short)(a[2] >> 16);
    alo[2] = (short)(a[3] & 0xffff); alo[3] = (short)(a[3] >> 16);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint16_t hi = bf16_bits(b[i]);
      bhi[i] = (short)hi;
      blo[i] = (short)bf16_bits(b[i] - bf16_value(hi));
    }
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, c, 0, 0, 0);
  }
};

template <int PREC, int ORDER, int KGB>
struct MixedShape {
  static constexpr int NP = np_of(ORDER);
  static constexpr int NFP = nfp_of(ORDER);
  static constexpr int NFN = kFaces * NFP;
  static constexpr int MB = (NP + 15) / 16;
  static constexpr int KG = MixedMma<PREC, KGB>::KG;
  static constexpr int KS = ((NP + 3) / 4 + KG - 1) / KG * KG;  // node k-steps per lane group
  static constexpr int KSL = (NFP + KG - 1) / KG * KG;         // LIFT k-steps per face
  static constexpr int AW = MixedMma<PREC, KGB>::A_WORDS;
  // A tiles, 32-bit words: [r][b][ks / KG][lane][AW] and [b][ks / KG][lane][AW]
  static constexpr int OP_WORDS = 3 * MB * (KS / KG) * 64 * AW;
  static constexpr int LIFT_WORDS = MB * (KSL / KG) * 64 * AW;
};

constexpr int kMixedFluxRow = 65;

// The next three choices hold off the direct flux, i.e. for BF16x3 but
// wave P3: every FP32 kernel runs the direct flux (mixed_direct_flux_of),
// which always stages the tile and holds its face geometry.  (Before it
// did, FP32 gained from each of them as well.)
//
// Whether the volume phase keeps the tile's FP32 state in LDS (NF * 16 * Np
// floats per wave) for the flux producer's own traces, as the FP64 MFMA
// kernel does.  On where it fits without lowering the blocks per CU that the
// registers allow and measured faster on MI355X: BF16x3 wave P1 -3%, Euler
// P1-P3 -2% to -5% (BF16x3 Maxwell P1 and P3 were 13% and 5% slower with
// it).
template <int EQ, int PREC, int ORDER>
constexpr bool mixed_stage_own_traces() {
  return (EQ == DG_EQ_WAVE && ORDER == 1) || (EQ == DG_EQ_EULER && ORDER <= 3);
}

// Whether the flux producer gathers without a branch (a boundary node loads
// its own trace and takes the boundary state by selection), so that all items
// are one basic block and their loads are in flight together, as in the FP64
// MFMA kernel.  Measured on MI355X, bit-identical results: BF16x3 wave P1
// -7%, P4 -15%, Euler P1-P4 -2% to -10%, Maxwell P1 -6%, P2 -2%.  Off where
// it was slower: BF16x3 Maxwell P3 (+25%) and P4 (+17%).
template <int EQ, int PREC, int ORDER>
constexpr bool mixed_branch_free_gathers() {
  return !(EQ == DG_EQ_MAXWELL && ORDER >= 3);
}

// Whether the kernel reads rx, the face geometry and the exterior map from
// the FP64 MFMA kernel's tile-ordered tables, producing the flux in that
// kernel's chunk order (one contiguous load per slot), instead of from the
// element-ordered geometry and map (on the direct flux, which has its own
// map and face geometry, it picks where rx comes from).  Measured on MI355X
// off the direct flux, bit-identical
// results: FP32 wave P1 -11%, P2 -3%, Euler P4 -5%, Maxwell P1 -12%, P2 -3%;
// BF16x3 wave P2 -5%, P3 -11%, Euler P2-P4 -3% to -11%, Maxwell P1 -12%,
// P2 -19%.  Off where it was even or slower: FP32 Euler P3 +61%, BF16x3
// Maxwell P3 +15%, P4 +19%, and the rest within +-2% (FP32 wave P3-P4,
// Euler P1-P2, Maxwell P3-P4; BF16x3 wave P1, P4 +4-5%, Euler P1).
template <int EQ, int PREC, int ORDER>
constexpr bool mixed_tile_tables() {
  if constexpr (PREC == DG_PREC_FP32)
    return (EQ != DG_EQ_EULER && ORDER <= 2) || (EQ == DG_EQ_EULER && ORDER == 4);
  if constexpr (EQ == DG_EQ_WAVE) return ORDER == 2 || ORDER == 3;
  if constexpr (EQ == DG_EQ_EULER) return ORDER >= 2;
  return ORDER <= 2;
}

// Whether the stage epilogue issues a field's res and q loads before the
// tile goes through LDS (all in flight together) or loads them in its store
// loop.  Preloading costs 2 NSLOT registers per field; in-loop where that
// would cost a wave per SIMD (FP32 wave P1-P2, Euler P2-P3; BF16x3 Maxwell
// P2).
template <int EQ, int PREC, int ORDER>
constexpr bool mixed_stage_preload() {
  if constexpr (PREC == DG_PREC_FP32)
    return !((EQ == DG_EQ_WAVE && ORDER <= 2) || (EQ == DG_EQ_EULER && (ORDER == 2 || ORDER == 3)));
  return !(EQ == DG_EQ_MAXWELL && ORDER == 2);
}

// Whether the Maxwell volume phase contracts the curl's combination
// w_{o,r} = sum_{x,a} eps_oxa rx[x][r] F_a (as the FP64 kernel does) rather
// than each field's reference derivatives, recombined after.  Measured on
// MI355X: FP32 P1 -12%, P2 -19%, P3 -11%, P4 -29%; BF16x3 P3 -6%, but P1
// +10% and P4 +6%, where splitting 18 combined B operands per node instead
// of 6 state values costs more than the recombination it saves.
template <int PREC, int ORDER>
constexpr bool mixed_maxwell_bform() {
  return PREC == DG_PREC_FP32 || ORDER == 3;
}

// Whether the FP32 wave kernel takes the FP64 kernel's direct flux with the
// velocity flux LIFTed in reference normals and in-tile neighbours from the
// staged state (rhs_mfma.h): its LIFT tiles in the direct k order, the
// exterior map from ext_d.  FP32 wave takes it at P1 too, where the FP64
// kernel does not (-14% on 40^3 x 6 tets, -15% on 56^3 x 6, -22% on 70^3 x 6,
// with the tile's face geometry and map held; +28% to even without).  FP32
// Euler and Maxwell take it at every order, with the flux LIFTed per field.
// Where the FP64 kernel does not run the direct flux (wave P1, Euler P1-P3,
// Maxwell P1, P3) the plan builds the direct map, with its in-tile entries,
// for this kernel alone.  Measured on MI355X:
// Maxwell P1 -4% (56^3 x 6 tets, 128 against 168 registers), P2 -8% (41^3 x
// 6) and -9% (30^3 x 6; 136 against 172 registers, so three waves per SIMD
// instead of two), P3 -20% (25^3 and 33^3 x 6) and -25% (45^3 x 6; 228
// against 320 registers, two waves per SIMD instead of one), P4 -9% on 384k
// tets and even on 94k; Euler P1 -6% (56^3 x 6 tets) and -4% (40^3 x 6; 128
// against 164 registers), P2 -1.5%, P3 -16% (33^3 x 6) and -19% (45^3 x 6;
// 236 against 276 registers, two waves per SIMD instead of one), P4 -3%.
//
// BF16x3 (four k per lane per instruction) takes it at P3, where the k order
// is packed: instruction kq, lane group g and its i-th k hold face-node pair
// 16 kq + 4 g + i, which is slot (4 kq + g, lane 16 i + n) of the exterior
// map ext_d.  Measured on MI355X: wave P3 -9%, Euler P3 -9%, Maxwell P3
// -38% (33^3 x 6 tets) and -30% (45^3 x 6; 252 against 292 registers, 63
// against 93 KB of LDS, two waves per SIMD instead of one); at P1-P2 Euler
// +20-24% and Maxwell +35-132%, where its registers cost waves.
__host__ __device__ constexpr bool mixed_direct_flux_of(int eq, int prec, int order) {
  // every FP32 kernel; BF16x3 at P3 (wave P2 measured 30% slower on it),
  // whose k order is packed with four k per lane
  return prec == DG_PREC_FP32 ||
         (prec == DG_PREC_BF16X3 && order == 3 && packed_direct_of(eq, order) &&
          bf16_kgroup_of(eq, order) == 4);
}
// Instructions of the direct LIFT contraction and the face-node pair of k
// index i of lane group g in instruction kq (-1: padding).
__host__ __device__ constexpr int mixed_direct_kq_of(int eq, int prec, int order) {
  return prec == DG_PREC_FP32 ? direct_ksteps_of(eq, order) : (4 * nfp_of(order) + 15) / 16;
}
__host__ __device__ constexpr int mixed_pair_of(int eq, int prec, int order, int kq, int g, int i) {
  return prec == DG_PREC_FP32 ? direct_pair_of(eq, order, kq, g)
                              : (16 * kq + 4 * g + i < 4 * nfp_of(order) ? 16 * kq + 4 * g + i : -1);
}

// Whether the flux producer off the direct flux reads its face geometry
// from a per-wave LDS copy of the tile's, formed from rx (Fscale n = rx^T
// c_f, where the plan finds it on every face; else copied from the table)
// instead of two global loads per flux item (off the direct flux: BF16x3).
// Measured on MI355X, the cases that gained: BF16x3 Maxwell P1 -17%, wave
// P2 -7.5%, Maxwell P4 -6.6%, Euler P4 -3.7%, Maxwell P2 -2%.  Elsewhere up
// to +14% (BF16x3 wave P1), where its LDS or registers cost a wave.
template <int EQ, int PREC, int ORDER>
constexpr bool mixed_face_geo_lds() {
  return (EQ == DG_EQ_WAVE && ORDER == 2) || (EQ == DG_EQ_EULER && ORDER == 4) ||
         (EQ == DG_EQ_MAXWELL && ORDER != 3);
}

// STAGE: instead of storing the FP32 RHS, the kernel applies the stage
// res = a res + dt rhs, q = q + b res, q32 = (float)q to it in its output
// phase (the FP32 RHS tile is the one the unfused path would store).
template <int EQ, int PREC, int ORDER, bool STAGE = false>
__global__ __launch_bounds__(kBlock) void rhs_mixed_kernel(MixedArgs a) {
  constexpr int KGB = bf16_kgroup_of(EQ, ORDER);
  constexpr bool MXB = mixed_maxwell_bform<PREC, ORDER>();
  using S = MixedShape<PREC, ORDER, KGB>;
  using M = MixedMma<PREC, KGB, bf16_hw_split_of(EQ, ORDER)>;
  using Flux = typename FluxOf<EQ>::type;
  constexpr int NP = S::NP, NFP = S::NFP, NFN = S::NFN, MB = S::MB;
  constexpr int KS = S::KS, KSL = S::KSL, KG = S::KG, AW = S::AW;
  constexpr int NF = nfields_of(EQ);
  constexpr int WAVES = kBlock / 64;
  constexpr int FROW = kMixedFluxRow;
  // direct flux (FP32 wave): LIFT k-steps and tiles in the direct order, no
  // flux buffer
  constexpr bool DFM = mixed_direct_flux_of(EQ, PREC, ORDER);
  constexpr int KSD = direct_ksteps_of(EQ, ORDER);         // the FP64 map's k-steps
  constexpr int KQD = mixed_direct_kq_of(EQ, PREC, ORDER); // direct LIFT instructions
  constexpr int LW = DFM ? MB * KQD * 64 * AW : S::LIFT_WORDS;
  // flux buffer: KSL k-steps x NF fields x FROW; output: 16 x NP of one field
  constexpr int WB = (!DFM && KSL * NF * FROW > 16 * NP) ? KSL * NF * FROW : 16 * NP;

  __shared__ uint32_t s_op[S::OP_WORDS];
  __shared__ uint32_t s_lift[LW];
  // direct flux at P3: the tile's face geometry [face * 16 + n][4] in LDS and
  // every k-step's map entry loaded at the tile's start (measured: FP32 wave
  // P3 -10%; at P2 +5% and at P4 even, where its LDS costs a block per CU)
  // the tile's face geometry and exterior map held for the whole tile (DFG):
  // P3, and FP32 P1-P2 with the geometry formed from rx (FP32 P2 -11.5%, P1
  // -14% to -22%; FP32 wave P4 +16%), and FP32 Euler and Maxwell (Maxwell
  // P4: -13% to -21% against the geometry loaded per k-step)
  constexpr bool DFG = DFM && (ORDER == 3 || (ORDER <= 2 && PREC == DG_PREC_FP32) ||
                               EQ != DG_EQ_WAVE);
  constexpr bool FGM = !DFM && mixed_face_geo_lds<EQ, PREC, ORDER>();
  constexpr bool FGL = DFG || FGM;   // the tile's face geometry in LDS
  __shared__ double2 s_fgeo[kBlock / 64][FGL ? 128 : 1];
  __shared__ int s_fmask[NFN];
  __shared__ float s_buf[WAVES][WB];
  // (the direct flux has no flux buffer and always stages the tile)
  constexpr bool ST = mixed_stage_own_traces<EQ, PREC, ORDER>() || DFM;
  __shared__ float s_stage[WAVES][ST ? NF * 16 * NP : 1];
  float* const stg = s_stage[threadIdx.x >> 6];

  const int tid = threadIdx.x;
  {
    const uint32_t* op = (const uint32_t*)a.op_tiles;
    const uint32_t* lf = (const uint32_t*)a.lift_tiles;
    for (int k = tid; k < S::OP_WORDS; k += kBlock) s_op[k] = op[k];
    for (int k = tid; k < LW; k += kBlock) s_lift[k] = lf[k];
    for (int k = tid; k < NFN; k += kBlock) s_fmask[k] = a.fmask[k];
  }
  __syncthreads();

  const int lane = tid & 63;
  const int g = lane >> 4;
  const int n = lane & 15;
  const int slot = block_in_xcd() * WAVES + (tid >> 6);
  const int nslots = blocks_per_xcd() * WAVES;
  const int32_t p_end = a.sched.part[block_xcd() + 1];
  const double c0 = a.p0;
  float* const buf = s_buf[tid >> 6];

  auto a_op = [&](int r, int b, int kq) {
    return &s_op[((((r * MB + b) * (KS / KG)) + kq) * 64 + lane) * AW];
  };
  auto a_lift = [&](int b, int kq) {
    return &s_lift[(((b * (KSL / KG)) + kq) * 64 + lane) * AW];
  };

  // the next tile's index is loaded while this tile runs (sched_at: through
  // the scalar cache)
  int32_t pos = __builtin_amdgcn_readfirstlane(a.sched.part[block_xcd()] + slot);
  int32_t next_tile = pos < p_end ? sched_at(a.sched.order, pos) : 0;
  for (; pos < p_end; pos += nslots) {
    const int64_t tile = next_tile;
    if (pos + nslots < p_end) next_tile = sched_at(a.sched.order, pos + nslots);
    // direct flux: the tile's face geometry into LDS and every k-step's map
    // entry, with the volume operands
    double2* const fgl = s_fgeo[tid >> 6];
    constexpr bool MAPA = DFG && PREC == DG_PREC_FP32;   // map entries preloaded
    int32_t ext_all[MAPA ? KSD : 1];
    // face geometry in LDS: on the direct flux, BF16x3 wave copies the
    // tile's table (forming it from rx measured +1.7% there) and FP32 wave
    // forms it from rx below (the plan refuses that kernel on other
    // geometry); FP32 Euler and Maxwell on the direct flux and the kernels off it
    // (FGM) form it from rx where the plan found Fscale n = rx^T c_f on
    // every face, else copy it
    constexpr bool DFG_RX = DFG && PREC == DG_PREC_FP32 && EQ == DG_EQ_WAVE;
    const bool geo_rx = DFG_RX || ((FGM || (DFG && EQ != DG_EQ_WAVE)) && a.face_geo_rx != 0);
    if constexpr (FGL) {
      if (!geo_rx) {
        const double2* src = reinterpret_cast<const double2*>(a.fgeo) + tile * 128;
        const double2 f0 = src[lane], f1 = src[lane + 64];
        fgl[lane] = f0;
        fgl[lane + 64] = f1;
      }
    }
    if constexpr (DFG) {
      if constexpr (MAPA)
#pragma unroll
        for (int ks = 0; ks < KSD; ++ks) ext_all[ks] = a.ext_d[(tile * KSD + ks) * 64 + lane];
    }
    const int64_t gel = tile * 16 + n;
    const bool valid = gel < a.nelem;
    const int64_t el = valid ? gel : a.nelem - 1;   // clamped for loads
    constexpr bool TT = mixed_tile_tables<EQ, PREC, ORDER>();
    double rx[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) rx[k] = TT ? a.vgeo[(tile * 9 + k) * 16 + n] : a.geo[el * kGeo + k];
    if constexpr (FGL) if (geo_rx) {
      // the tile's face geometry from rx, Fscale n = rx^T c_g for face g of
      // element n (lane (g, n)), as the FP64 kernel forms it, on geometry
      // that satisfies it (geo_rx above).  Measured on MI355X: FP32 wave P3
      // -8% (162k tets), -13% (216k tets)
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

    f32x4 acc[NF][MB];
#pragma unroll
    for (int c = 0; c < NF; ++c)
#pragma unroll
      for (int b = 0; b < MB; ++b) acc[c][b] = f32x4{0.f, 0.f, 0.f, 0.f};

    // {{{ volume terms (B operands built in FP64, rounded once to FP32)
    // (k-step loops unrolled by 2; measured on MI355X against no unrolling:
    // FP32 wave P3 -9%, wave P4 -7.5%, Maxwell P3 -9%, Maxwell P4 -3.5%,
    // Euler -3% to even; BF16x3 -3% to +1.6%)

#pragma unroll 2
    for (int kq = 0; kq < KS / KG; ++kq) {
      if constexpr (EQ == DG_EQ_WAVE) {
        float ub[KG], wb[3][KG];
#pragma unroll
        for (int i = 0; i < KG; ++i) {
          const int j = g * KS + kq * KG + i;
          double u = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
          if (j < NP) {
            const int64_t idx = el * NP + j;
            u = a.q[0][idx];
            vx = a.q[1][idx];
            vy = a.q[2][idx];
            vz = a.q[3][idx];
            if constexpr (ST) {
              stg[0 * 16 * NP + n * NP + j] = (float)u;
              stg[1 * 16 * NP + n * NP + j] = (float)vx;
              stg[2 * 16 * NP + n * NP + j] = (float)vy;
              stg[3 * 16 * NP + n * NP + j] = (float)vz;
            }
          }
          ub[i] = (float)u;
#pragma unroll
          for (int r = 0; r < 3; ++r)
            wb[r][i] = (float)(rx[r] * vx + rx[3 + r] * vy + rx[6 + r] * vz);
        }
        // t_r = D_r u lives in acc[1 + r] until the combination below
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const uint32_t* av = a_op(r, b, kq);
            uint32_t aw[AW];
#pragma unroll
            for (int w = 0; w < AW; ++w) aw[w] = av[w];
            acc[1 + r][b] = M::mma(aw, ub, acc[1 + r][b]);
            acc[0][b] = M::mma(aw, wb[r], acc[0][b]);
          }
      } else if constexpr (EQ == DG_EQ_EULER) {
        const double gamma = c0;
        float wb[5][3][KG];
#pragma unroll
        for (int i = 0; i < KG; ++i) {
          const int j = g * KS + kq * KG + i;
          if (j < NP) {
            const int64_t idx = el * NP + j;
            const double rho = a.q[0][idx], e = a.q[1][idx];
            const double m[3] = {a.q[2][idx], a.q[3][idx], a.q[4][idx]};
            if constexpr (ST) {
              stg[0 * 16 * NP + n * NP + j] = (float)rho;
              stg[1 * 16 * NP + n * NP + j] = (float)e;
#pragma unroll
              for (int x = 0; x < 3; ++x) stg[(2 + x) * 16 * NP + n * NP + j] = (float)m[x];
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
                wb[c][r][i] = (float)(rx[r] * f[c][0] + rx[3 + r] * f[c][1] + rx[6 + r] * f[c][2]);
          } else {
#pragma unroll
            for (int c = 0; c < 5; ++c) wb[c][0][i] = wb[c][1][i] = wb[c][2][i] = 0.f;
          }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const uint32_t* av = a_op(r, b, kq);
            uint32_t aw[AW];
#pragma unroll
            for (int w = 0; w < AW; ++w) aw[w] = av[w];
#pragma unroll
            for (int c = 0; c < 5; ++c) acc[c][b] = M::mma(aw, wb[c][r], acc[c][b]);
          }
      } else if constexpr (MXB) {
        // Maxwell: E_t = curl H, H_t = -curl E with the curl's rx /
        // Levi-Civita combination in B-operand space, as in the FP64 kernel:
        // (curl F)_o = sum_r D_r w_{o,r}, w_{o,r} = sum_{x,a} eps_oxa rx[x][r] F_a
        double f[6][KG];
#pragma unroll
        for (int i = 0; i < KG; ++i) {
          const int j = g * KS + kq * KG + i;
#pragma unroll
          for (int c = 0; c < 6; ++c) f[c][i] = 0.0;
          if (j < NP) {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
              const float v = a.q[c][el * NP + j];
              f[c][i] = v;
              if constexpr (ST) stg[c * 16 * NP + n * NP + j] = v;
            }
          }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          float wb[6][KG];
#pragma unroll
          for (int i = 0; i < KG; ++i)
#pragma unroll
            for (int o = 0; o < 3; ++o) {
              double ch = 0.0, ce = 0.0;
#pragma unroll
              for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                  const int sg = levi(o, x, c);
                  if (sg == 0) continue;
                  ch += (double)sg * rx[3 * x + r] * f[3 + c][i];
                  ce += (double)sg * rx[3 * x + r] * f[c][i];
                }
              wb[o][i] = (float)ch;
              wb[3 + o][i] = (float)(-ce);
            }
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const uint32_t* av = a_op(r, b, kq);
            uint32_t aw[AW];
#pragma unroll
            for (int w = 0; w < AW; ++w) aw[w] = av[w];
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c][b] = M::mma(aw, wb[c], acc[c][b]);
          }
        }
      }
    }
    if constexpr (EQ == DG_EQ_WAVE) {
#pragma unroll
      for (int b = 0; b < MB; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double t0 = acc[1][b][v], t1 = acc[2][b][v], t2 = acc[3][b][v];
          acc[0][b][v] = (float)(-c0 * (double)acc[0][b][v]);
          if constexpr (DFM) {
            // t_r = -c D_r u stay, to take the velocity flux (combined below)
            acc[1][b][v] = (float)(-c0 * t0);
            acc[2][b][v] = (float)(-c0 * t1);
            acc[3][b][v] = (float)(-c0 * t2);
          } else {
#pragma unroll
            for (int x = 0; x < 3; ++x)
              acc[1 + x][b][v] = (float)(-c0 * (rx[3 * x] * t0 + rx[3 * x + 1] * t1 + rx[3 * x + 2] * t2));
          }
        }
    }
    if constexpr (EQ == DG_EQ_MAXWELL && !MXB) {
      // E_t = curl H, H_t = -curl E from t_{c,r} = D_r F_c, one field at a time
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        f32x4 tr[3][MB];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int b = 0; b < MB; ++b) tr[r][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
        for (int kq = 0; kq < KS / KG; ++kq) {
          float ub[KG];
#pragma unroll
          for (int i = 0; i < KG; ++i) {
            const int j = g * KS + kq * KG + i;
            ub[i] = j < NP ? a.q[c][el * NP + j] : 0.f;
            if constexpr (ST)
              if (j < NP) stg[c * 16 * NP + n * NP + j] = ub[i];
          }
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int b = 0; b < MB; ++b) {
              const uint32_t* av = a_op(r, b, kq);
              uint32_t aw[AW];
#pragma unroll
              for (int w = 0; w < AW; ++w) aw[w] = av[w];
              tr[r][b] = M::mma(aw, ub, tr[r][b]);
            }
        }
        const int comp = c % 3;
        const bool is_h = c >= 3;
#pragma unroll
        for (int b = 0; b < MB; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const double t0 = tr[0][b][v], t1 = tr[1][b][v], t2 = tr[2][b][v];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
              const double d = rx[3 * x] * t0 + rx[3 * x + 1] * t1 + rx[3 * x + 2] * t2;
#pragma unroll
              for (int o = 0; o < 3; ++o) {
                const int sgn = (o == x || x == comp || o == comp) ? 0
                                : (((x - o + 3) % 3 == 1) ? 1 : -1);
                if (sgn == 0) continue;
                const int out = is_h ? o : 3 + o;
                acc[out][b][v] += (float)((is_h ? sgn : -sgn) * d);
              }
            }
          }
      }
    }

    // }}}

    // {{{ surface terms: flux in FP64, staged as FP32 B operands

    if constexpr (ST || FGL) asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");   // staged state / face geometry complete

    if constexpr (DFM) {
      // FP32 wave, the FP64 kernel's direct flux: at k-step ks lane (g, n)
      // evaluates the upwind flux at face-node pair direct_pair_of(ks, g) of
      // element n; the u part goes to acc[0], the velocity part s n to the
      // t_r in reference normals (c_{f,r}); in-tile neighbours from the
      // staged state, the others by buffer loads whose out-of-range lanes
      // touch no memory
      static_assert(ST, "the direct flux reads the staged tile");
      const uint32_t field_bytes = (uint32_t)(a.nelem * NP * 4);
      __amdgpu_buffer_rsrc_t rsq[NF];
#pragma unroll
      for (int c = 0; c < NF; ++c)
        rsq[c] = __builtin_amdgcn_make_buffer_rsrc((void*)a.q[c], (short)0, field_bytes, kBufferWord3);
      if constexpr (EQ != DG_EQ_WAVE) {
        // Euler and Maxwell: a lane evaluates the FP64 flux at each
        // face-node pair that feeds its LIFT B operand (FP32: pair
        // direct_pair_of(ks, g) of element n at k-step ks; BF16x3: its KG
        // pairs per instruction) and feeds every field's Fscale-scaled flux
        // to it; the boundary state is Flux::boundary of the lane's own
        // trace (for Maxwell's PEC walls, the trace with E negated) (the
        // loops rolled: the form measured)
        if constexpr (PREC != DG_PREC_FP32) {
          // BF16x3 (P3): per instruction, its KG face-node pairs per lane,
          // each pair's map entry from its packed slot of ext_d
#pragma unroll 1
          for (int kq = 0; kq < KQD; ++kq) {
            float bv[NF][KG];
#pragma unroll
            for (int i = 0; i < KG; ++i) {
              const int fp = mixed_pair_of(EQ, PREC, ORDER, kq, g, i);
              const bool fok = fp >= 0;
              const int pp = fok ? fp : 0;
              const int f = pp / NFP;
              const int vloc = s_fmask[pp];
              const double2 g01 = fgl[2 * (f * 16 + n)];
              const double2 g23 = fgl[2 * (f * 16 + n) + 1];
              const int32_t ext = a.ext_d[(tile * KSD + pp / 4) * 64 + (pp % 4) * 16 + n];
              const bool loc = ext >= kInTile;
              const bool glob = ext >= 0 && !loc;
              const int lofs = loc ? ((ext >> 8) & 15) * NP + (ext & 255) : n * NP + vloc;
              const uint32_t boff = glob ? (uint32_t)ext * 4u : field_bytes;
              const float wl = glob ? 0.f : 1.f;
              double qm[NF], qp[NF], out[NF];
#pragma unroll
              for (int c = 0; c < NF; ++c) {
                qm[c] = stg[c * 16 * NP + n * NP + vloc];
                const float gv = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsq[c], boff, 0, 0));
                qp[c] = (double)__builtin_fmaf(stg[c * 16 * NP + lofs], wl, gv);
              }
              const double nrm[3] = {g01.x, g01.y, g23.x};
              if constexpr (Flux::mirror_sign(0) != 0) {
#pragma unroll
                for (int c = 0; c < NF; ++c)
                  if (Flux::mirror_sign(c) < 0) qp[c] = ext >= 0 ? qp[c] : -qp[c];
              } else {
                double qb[NF];
                Flux::boundary(qm, nrm, qb);
#pragma unroll
                for (int c = 0; c < NF; ++c) qp[c] = ext >= 0 ? qp[c] : qb[c];
              }
              Flux::flux(qm, qp, nrm, c0, out);
#pragma unroll
              for (int c = 0; c < NF; ++c) bv[c][i] = fok ? (float)(g23.y * out[c]) : 0.f;
            }
#pragma unroll
            for (int b = 0; b < MB; ++b) {
              const uint32_t* av = &s_lift[((b * KQD + kq) * 64 + lane) * AW];
              uint32_t aw[AW];
#pragma unroll
              for (int w = 0; w < AW; ++w) aw[w] = av[w];
#pragma unroll
              for (int c = 0; c < NF; ++c) acc[c][b] = M::mma(aw, bv[c], acc[c][b]);
            }
          }
        } else {
#pragma unroll 1
        for (int ks = 0; ks < KSD; ++ks) {
          const int fp = direct_pair_of(EQ, ORDER, ks, g);
          const bool fok = fp >= 0;
          const int f = fok ? fp / NFP : ks / ((NFP + 3) / 4);
          const int vloc = s_fmask[fok ? fp : f * NFP];
          static_assert(DFG && MAPA, "FP32 Euler and Maxwell hold the tile's geometry and map");
          const double2 g01 = fgl[2 * (f * 16 + n)];
          const double2 g23 = fgl[2 * (f * 16 + n) + 1];
          const int32_t ext = ext_all[ks];
          const bool loc = ext >= kInTile;
          const bool glob = ext >= 0 && !loc;
          const int lofs = loc ? ((ext >> 8) & 15) * NP + (ext & 255) : n * NP + vloc;
          const uint32_t boff = glob ? (uint32_t)ext * 4u : field_bytes;
          const float wl = glob ? 0.f : 1.f;
          double qm[NF], qp[NF], out[NF];
#pragma unroll
          for (int c = 0; c < NF; ++c) {
            qm[c] = stg[c * 16 * NP + n * NP + vloc];
            const float gv = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsq[c], boff, 0, 0));
            qp[c] = (double)__builtin_fmaf(stg[c * 16 * NP + lofs], wl, gv);
          }
          const double nrm[3] = {g01.x, g01.y, g23.x};
          if constexpr (Flux::mirror_sign(0) != 0) {
#pragma unroll
            for (int c = 0; c < NF; ++c)
              if (Flux::mirror_sign(c) < 0) qp[c] = ext >= 0 ? qp[c] : -qp[c];
          } else {
            double qb[NF];
            Flux::boundary(qm, nrm, qb);
#pragma unroll
            for (int c = 0; c < NF; ++c) qp[c] = ext >= 0 ? qp[c] : qb[c];
          }
          Flux::flux(qm, qp, nrm, c0, out);
          float bv[NF][1];
#pragma unroll
          for (int c = 0; c < NF; ++c) bv[c][0] = fok ? (float)(g23.y * out[c]) : 0.f;
#pragma unroll
          for (int b = 0; b < MB; ++b) {
            const uint32_t* av = &s_lift[(b * KSD + ks) * 64 + lane];
#pragma unroll
            for (int c = 0; c < NF; ++c) acc[c][b] = M::mma(av, bv[c], acc[c][b]);
          }
        }
        }
      } else {
      // Two forms of the same loop: per k-step with one k per lane (FP32 P2,
      // P3), and per instruction over its KG k (BF16x3, and FP32 P4).  The
      // compiler schedules them differently: measured on MI355X, the first
      // is 24-39% faster at FP32 P2-P3 (the second serialises the buffer
      // loads there), the second 14% faster at FP32 P4.
      if constexpr (PREC == DG_PREC_FP32 && ORDER <= 3) {
#pragma unroll
      for (int ks = 0; ks < KSD; ++ks) {
        const int fp = direct_pair_of(EQ, ORDER, ks, g);
        const bool fok = fp >= 0;
        const int f = fok ? fp / NFP : ks / ((NFP + 3) / 4);
        const int vloc = s_fmask[fok ? fp : f * NFP];
        double2 g01, g23;
        int32_t ext;
        if constexpr (DFG) {
          g01 = fgl[2 * (f * 16 + n)];
          g23 = fgl[2 * (f * 16 + n) + 1];
          ext = ext_all[MAPA ? ks : 0];
        } else {
          const double2* pgeo = reinterpret_cast<const double2*>(a.fgeo + (tile * 64 + f * 16 + n) * 4);
          g01 = pgeo[0];
          g23 = pgeo[1];
          ext = a.ext_d[(tile * KSD + ks) * 64 + lane];
        }
        const bool loc = ext >= kInTile;
        const bool glob = ext >= 0 && !loc;
        const int lofs = loc ? ((ext >> 8) & 15) * NP + (ext & 255) : n * NP + vloc;
        const uint32_t boff = glob ? (uint32_t)ext * 4u : field_bytes;
        const float wl = glob ? 0.f : 1.f;
        double qm[4], qp[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          qm[c] = stg[c * 16 * NP + n * NP + vloc];
          const float gv = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsq[c], boff, 0, 0));
          qp[c] = (double)__builtin_fmaf(stg[c * 16 * NP + lofs], wl, gv);
        }
        qp[0] = ext >= 0 ? qp[0] : -qp[0];     // mirror boundary: u+ = -u-, v+ = v-
        const double am = qm[1] * g01.x + qm[2] * g01.y + qm[3] * g23.x;
        const double bp = qp[1] * g01.x + qp[2] * g01.y + qp[3] * g23.x;
        const double h = 0.5 * c0;
        float bu[1], bs[1];
        bu[0] = fok ? (float)((h * g23.y) * ((am + bp) + (qp[0] - qm[0]))) : 0.f;
        const double sv = fok ? h * ((qm[0] + qp[0]) + (bp - am)) : 0.0;
        // faces this k-step covers, over its four lane groups
        const int f_lo = direct_pair_of(EQ, ORDER, ks, 0) >= 0 ? direct_pair_of(EQ, ORDER, ks, 0) / NFP
                                                                : ks / ((NFP + 3) / 4);
        int f_hi = f_lo;
#pragma unroll
        for (int gg = 1; gg < 4; ++gg)
          if (direct_pair_of(EQ, ORDER, ks, gg) >= 0) f_hi = direct_pair_of(EQ, ORDER, ks, gg) / NFP;
#pragma unroll
        for (int b = 0; b < MB; ++b) {
          const uint32_t* av = &s_lift[(b * KSD + ks) * 64 + lane];
          acc[0][b] = M::mma(av, bu, acc[0][b]);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            bool any = false;
#pragma unroll
            for (int ff = 0; ff < 4; ++ff) any |= ff >= f_lo && ff <= f_hi && ref_face_coef(ff, r) != 0;
            if (any) {
              bs[0] = (float)((double)ref_face_coef(f, r) * sv);
              acc[1 + r][b] = M::mma(av, bs, acc[1 + r][b]);
            }
          }
        }
      }
      } else {
#pragma unroll
      for (int kq = 0; kq < KQD; ++kq) {
        float bu[KG], sv[KG];
        int fi[KG];
#pragma unroll
        for (int i = 0; i < KG; ++i) {
          const int fp = mixed_pair_of(EQ, PREC, ORDER, kq, g, i);
          const bool fok = fp >= 0;
          const int f = fok ? fp / NFP : (PREC == DG_PREC_FP32 ? kq / ((NFP + 3) / 4) : 0);
          fi[i] = f;
          const int vloc = s_fmask[fok ? fp : f * NFP];
          double2 g01, g23;
          int32_t ext;
          if constexpr (DFG) {
            g01 = fgl[2 * (f * 16 + n)];
            g23 = fgl[2 * (f * 16 + n) + 1];
          } else {
            const double2* pgeo = reinterpret_cast<const double2*>(a.fgeo + (tile * 64 + f * 16 + n) * 4);
            g01 = pgeo[0];
            g23 = pgeo[1];
          }
          if constexpr (PREC == DG_PREC_FP32) {
            if constexpr (MAPA) ext = ext_all[kq];
            else ext = a.ext_d[(tile * KSD + kq) * 64 + lane];
          } else {
            // pair fp sits at slot (fp / 4, lane 16 (fp % 4) + n) of ext_d;
            // padding reads pair 0's (valid) and contributes nothing
            const int pp = fok ? fp : 0;
            ext = a.ext_d[(tile * KSD + pp / 4) * 64 + (pp % 4) * 16 + n];
          }
          const bool loc = ext >= kInTile;
          const bool glob = ext >= 0 && !loc;
          const int lofs = loc ? ((ext >> 8) & 15) * NP + (ext & 255) : n * NP + vloc;
          const uint32_t boff = glob ? (uint32_t)ext * 4u : field_bytes;
          const float wl = glob ? 0.f : 1.f;
          double qm[4], qp[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            qm[c] = stg[c * 16 * NP + n * NP + vloc];
            const float gv = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsq[c], boff, 0, 0));
            qp[c] = (double)__builtin_fmaf(stg[c * 16 * NP + lofs], wl, gv);
          }
          qp[0] = ext >= 0 ? qp[0] : -qp[0];     // mirror boundary: u+ = -u-, v+ = v-
          const double am = qm[1] * g01.x + qm[2] * g01.y + qm[3] * g23.x;
          const double bp = qp[1] * g01.x + qp[2] * g01.y + qp[3] * g23.x;
          const double h = 0.5 * c0;
          bu[i] = fok ? (float)((h * g23.y) * ((am + bp) + (qp[0] - qm[0]))) : 0.f;
          sv[i] = fok ? (float)(h * ((qm[0] + qp[0]) + (bp - am))) : 0.f;
        }
        // faces this instruction covers, over its lane groups and k
        int f_lo = 3, f_hi = 0;
#pragma unroll
        for (int gg = 0; gg < 4; ++gg)
#pragma unroll
          for (int i = 0; i < KG; ++i) {
            const int pq = mixed_pair_of(EQ, PREC, ORDER, kq, gg, i);
            if (pq >= 0) {
              f_lo = pq / NFP < f_lo ? pq / NFP : f_lo;
              f_hi = pq / NFP > f_hi ? pq / NFP : f_hi;
            }
          }
#pragma unroll
        for (int b = 0; b < MB; ++b) {
          const uint32_t* av = &s_lift[((b * KQD + kq) * 64 + lane) * AW];
          uint32_t aw[AW];
#pragma unroll
          for (int w = 0; w < AW; ++w) aw[w] = av[w];
          acc[0][b] = M::mma(aw, bu, acc[0][b]);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            bool any = false;
#pragma unroll
            for (int ff = 0; ff < 4; ++ff) any |= ff >= f_lo && ff <= f_hi && ref_face_coef(ff, r) != 0;
            if (any) {
              float bs[KG];
#pragma unroll
              for (int i = 0; i < KG; ++i) bs[i] = (float)ref_face_coef(fi[i], r) * sv[i];
              acc[1 + r][b] = M::mma(aw, bs, acc[1 + r][b]);
            }
          }
        }
      }
      }
#pragma unroll
      for (int b = 0; b < MB; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double t0 = acc[1][b][v], t1 = acc[2][b][v], t2 = acc[3][b][v];
#pragma unroll
          for (int x = 0; x < 3; ++x)
            acc[1 + x][b][v] = (float)(rx[3 * x] * t0 + rx[3 * x + 1] * t1 + rx[3 * x + 2] * t2);
        }
      }
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");   // reads of the staged tile done
    } else {

    // With the tile tables the items run in the FP64 MFMA kernel's chunk
    // order (C LIFT k-steps per chunk; C divides Nfp), so the face geometry
    // and the exterior map come with one contiguous load per slot; without,
    // all KSL k-steps are one chunk and the element-ordered tables are read.
    // BF16x3 pads the k-steps to KSL; the padding is zero.
    constexpr int C = TT ? flux_chunk_of(EQ, ORDER) : KSL;
    constexpr int NCH = TT ? NFP / C : 1;
    static_assert(!TT || NFP % C == 0, "whole flux chunks");
    if constexpr (TT) {
#pragma unroll
      for (int kk = NFP; kk < KSL; ++kk)
#pragma unroll
        for (int c = 0; c < NF; ++c) buf[(kk * NF + c) * FROW + lane] = 0.f;
    }
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
    for (int it = 0; it < C; ++it) {
      const int item = it * 64 + lane;
      const int kk = ch * C + item % C;
      const int p = item / C;                     // p = face * 16 + element
      const int fg = p >> 4, fe = p & 15;
      float outf[NF];
      if (kk < NFP) {
        int64_t pe = tile * 16 + fe;
        pe = pe < a.nelem ? pe : a.nelem - 1;
        double nrm[3], fscale;
        if constexpr (FGM) {
          const double2 g01 = fgl[2 * p], g23 = fgl[2 * p + 1];
          nrm[0] = g01.x;
          nrm[1] = g01.y;
          nrm[2] = g23.x;
          fscale = g23.y;
        } else if constexpr (TT) {
          const double2* pgeo = reinterpret_cast<const double2*>(a.fgeo + (tile * 64 + p) * 4);
          const double2 g01 = pgeo[0], g23 = pgeo[1];
          nrm[0] = g01.x;
          nrm[1] = g01.y;
          nrm[2] = g23.x;
          fscale = g23.y;
        } else {
          const double* pgeo = a.geo + pe * kGeo + 9 + 4 * fg;
          nrm[0] = pgeo[0];
          nrm[1] = pgeo[1];
          nrm[2] = pgeo[2];
          fscale = pgeo[3];
        }
        const int fn = fg * NFP + kk;
        const int64_t vn = pe * NP + s_fmask[fn];
        double qm[NF], qp[NF], out[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c)
          qm[c] = ST ? (double)stg[c * 16 * NP + fe * NP + s_fmask[fn]] : (double)a.q[c][vn];
        const int32_t ext = TT ? a.ext_t[tile * flux_items_of(EQ, ORDER) + (ch * C + it) * 64 + lane]
                               : a.emap[pe * NFN + fn];
        if constexpr (mixed_branch_free_gathers<EQ, PREC, ORDER>()) {
          const int64_t src = ext >= 0 ? (int64_t)ext : vn;
#pragma unroll
          for (int c = 0; c < NF; ++c) qp[c] = a.q[c][src];
          double qb[NF];
          Flux::boundary(qm, nrm, qb);
#pragma unroll
          for (int c = 0; c < NF; ++c) qp[c] = ext >= 0 ? qp[c] : qb[c];
        } else if (ext >= 0) {
#pragma unroll
          for (int c = 0; c < NF; ++c) qp[c] = a.q[c][ext];
        } else {
          Flux::boundary(qm, nrm, qp);
        }
        Flux::flux(qm, qp, nrm, c0, out);
#pragma unroll
        for (int c = 0; c < NF; ++c) outf[c] = (float)(fscale * out[c]);
      } else {
#pragma unroll
        for (int c = 0; c < NF; ++c) outf[c] = 0.f;
      }
#pragma unroll
      for (int c = 0; c < NF; ++c) buf[(kk * NF + c) * FROW + p] = outf[c];
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll 1
    for (int kq = 0; kq < KSL / KG; ++kq) {
#pragma unroll
      for (int b = 0; b < MB; ++b) {
        const uint32_t* av = a_lift(b, kq);
        uint32_t aw[AW];
#pragma unroll
        for (int w = 0; w < AW; ++w) aw[w] = av[w];
#pragma unroll
        for (int c = 0; c < NF; ++c) {
          float fb[KG];
#pragma unroll
          for (int i = 0; i < KG; ++i) fb[i] = buf[((kq * KG + i) * NF + c) * FROW + lane];
          acc[c][b] = M::mma(aw, fb, acc[c][b]);
        }
      }
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    }

    // }}}

    // {{{ output through the buffer, one field at a time

    // STAGE: each lane's slots of the output tile, idx = lane + 64 t
    constexpr int NSLOT = (16 * NP + 63) / 64;
    constexpr bool PRE = STAGE && mixed_stage_preload<EQ, PREC, ORDER>();
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      // STAGE: this field's res and q loads are issued before the tile goes
      // through LDS, so they are in flight together
      double st_res[PRE ? NSLOT : 1], st_q[PRE ? NSLOT : 1];
      if constexpr (PRE) {
        const int64_t off = tile * 16 * NP;
#pragma unroll
        for (int t = 0; t < NSLOT; ++t) {
          const int idx = lane + 64 * t;
          const bool ok = idx < 16 * NP && tile * 16 + idx / NP < a.nelem;
          // a == 0 (the first stage) does not read res, so it may start uninitialised
          st_res[t] = ok && a.st_a != 0.0 ? a.st_res[c][off + idx] : 0.0;
          st_q[t] = ok ? a.st_q[c][off + idx] : 0.0;
        }
      }
#pragma unroll
      for (int b = 0; b < MB; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * b + M::row(g, v);
          if (i < NP) buf[n * NP + i] = acc[c][b][v];
        }
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      if constexpr (STAGE) {
        const int64_t off = tile * 16 * NP;
#pragma unroll
        for (int t = 0; t < NSLOT; ++t) {
          const int idx = lane + 64 * t;
          if (idx < 16 * NP && tile * 16 + idx / NP < a.nelem) {
            const double r0 = PRE ? st_res[PRE ? t : 0] : (a.st_a != 0.0 ? a.st_res[c][off + idx] : 0.0);
            const double q0 = PRE ? st_q[PRE ? t : 0] : a.st_q[c][off + idx];
            const double r = (a.st_a != 0.0 ? a.st_a * r0 : 0.0) + a.st_dt * (double)buf[idx];
            const double qn = q0 + a.st_b * r;
            a.st_res[c][off + idx] = r;
            a.st_q[c][off + idx] = qn;
            a.st_q32[c][off + idx] = (float)qn;
          }
        }
      } else {
        // all of the lane's slots read from LDS, then stored
        float* dst = a.rhs[c] + tile * 16 * NP;
        float vals[NSLOT];
#pragma unroll
        for (int t = 0; t < NSLOT; ++t) {
          const int idx = lane + 64 * t;
          vals[t] = idx < 16 * NP ? buf[idx] : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < NSLOT; ++t) {
          const int idx = lane + 64 * t;
          if (idx < 16 * NP && tile * 16 + idx / NP < a.nelem)
            __builtin_nontemporal_store(vals[t], dst + idx);
        }
      }
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    }

    // }}}
  }
}

}  // namespace dg
