/*
 * dgfem.h -- C ABI of libdgfem.so, the gfx950 DG-FEM right-hand-side engine.
 *
 * The reference (kaushikcfd/actx_dgfem_suite) is pure Python: its hot path is
 * `compiled_rhs = actx.compile(rhs); compiled_rhs(*args)` (run/__main__.py:74-76,
 * :109, :126), where `rhs` is the wave / Euler / Maxwell operator built by
 * get_{wave,euler,maxwell}_rhs (equations/wave.py:239-285,
 * equations/euler.py:140-186, equations/maxwell.py:43-74) and the array
 * context turns the operator's einsum DAG into device kernels
 * (arraycontext/dgfem_optimizer.py:181-245).  This library is what a new array
 * context binds to in place of that code generator.  Each entry point below
 * names the reference interface it replaces.
 *
 * Conventions
 *  - All functions return 0 (DG_OK) on success or a positive dg_status code;
 *    dg_last_error_string() gives a message for the calling thread.
 *  - Field data are device pointers to C-contiguous (K, Np) float64 arrays, one
 *    per field, element-major: the layout of one meshmode DOFArray group.
 *    Field order, tetrahedra: wave (u, v_x, v_y, v_z); Euler (mass, energy,
 *    momentum x, y, z); Maxwell (E_x, E_y, E_z, H_x, H_y, H_z).  Triangles
 *    (dim = 2): wave (u, v_x, v_y); Euler (mass, energy, momentum x, y);
 *    Maxwell in TM form (E_z, H_x, H_y).
 *  - `fields_in` / `fields_out` are HOST arrays of nfields DEVICE pointers.
 *  - Launches are asynchronous on `stream`; nothing synchronises the host.
 *  - Hot-path calls allocate nothing.  Plans own their device constants.
 */
#ifndef DGFEM_H
#define DGFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dg_status {
  DG_OK = 0,
  DG_ERR_INVALID_ARGUMENT = 1,
  DG_ERR_UNSUPPORTED = 2,
  DG_ERR_HIP = 3,
  DG_ERR_OUT_OF_MEMORY = 4
} dg_status;

typedef enum dg_equation {
  DG_EQ_WAVE = 0,     /* equations/wave.py:121-200 */
  DG_EQ_EULER = 1,    /* equations/euler.py:168-179 (EulerOperator, LF, walls) */
  DG_EQ_MAXWELL = 2   /* equations/maxwell.py:57-59 (MaxwellOperator, PEC) */
} dg_equation;

typedef enum dg_kernel_variant {
  DG_KERNEL_AUTO = 0,   /* the measured-fastest per equation and order */
  DG_KERNEL_MFMA = 1,   /* fused kernel on v_mfma_f64_16x16x4f64 */
  DG_KERNEL_LDS = 2     /* one thread per node, LDS-staged (baseline) */
} dg_kernel_variant;

typedef enum dg_precision {
  DG_PREC_FP64 = 0,   /* the drop-in form: FP64 state, FP64 arithmetic */
  DG_PREC_FP32 = 1,   /* mixed: FP32 state, D/LIFT on FP32 MFMA, geometry/flux in FP64 */
  DG_PREC_BF16X3 = 2  /* mixed: as FP32 with D/LIFT as a 3-product bf16 split */
} dg_precision;

/* Everything a plan needs; all pointers are HOST memory and are copied.
 * Shapes are given for tetrahedra (d = 3 directions, F = 4 faces); triangles
 * (dim = 2) have d = 2, F = 3, Np = (N+1)(N+2)/2 and Nfp = N+1. */
typedef struct dg_plan_desc {
  int32_t equation;          /* dg_equation */
  int32_t order;             /* polynomial degree, 1..4 */
  int64_t nelements;         /* K */
  /* (d, Np, Np) row-major [r][i][j]: M^{-1} S_r^T, the weak derivative with
   * the inverse mass fused (mass_inverse_fuser.py:39-74) */
  const double* d_weak;
  /* (d, Np, Np): V_r V^{-1}, the strong reference derivative */
  const double* d_strong;
  /* (Np, F*Nfp): M^{-1} E, face mass fused with inverse mass
   * (mass_inverse_fuser.py:21-38) */
  const double* lift;
  /* (F, Nfp): volume node index of each face node */
  const int32_t* face_nodes;
  /* Tetrahedra (K, 25): rx[x][r] = dr_r/dx_x (9 values), then for faces
   * f=0..3: n_x, n_y, n_z, Fscale = sJ/J.  Triangles (K, 13): rx[x][r] (4
   * values), then for edges f=0..2: n_x, n_y, Fscale. */
  const double* geometry;
  /* (K, F*Nfp): flat index e2*Np + i2 of the exterior node of every face
   * node; boundary face nodes hold -(e*Np + i) - 1 (their own node, negated) */
  const int32_t* ext_map;
  /* wave: params[0] = c; Euler: params[0] = gamma; Maxwell: params[0] =
   * flux_type alpha (0 central .. 1 upwind), with epsilon = mu = 1 as in
   * equations/maxwell.py:57-59 */
  double params[4];
  int32_t variant;           /* dg_kernel_variant (triangles: AUTO or LDS) */
  int32_t dim;               /* 3 (or 0): tetrahedra; 2: triangles, which run
                              * the LDS-staged kernel with FP64 arithmetic */
  /* Needed by dg_rhs_vjp only (both NULL for a forward-only plan):
   * (Np, Np) reference inverse mass M^{-1}, and (Np, F*Nfp) reference face
   * mass E with LIFT = M^{-1} E (F = 4 faces, 3 edges for triangles) (mass_inverse_fuser.py:21-38 is the fusion
   * the adjoint splits apart again). */
  const double* mass_inv;
  const double* face_mass;
  /* Optional (NULL: storage order): (K, 3) element centroids.  The plan
   * uses them only to order its work for cache locality (each XCD sweeps a
   * compact part of the mesh); results do not depend on them. */
  const double* centroids;
} dg_plan_desc;

typedef struct dg_plan dg_plan;

/* Builds a plan: copies operators, geometry and the exterior map to the
 * current device.  Replaces the reference's compile-time constant folding of
 * geometry and gather indices (arraycontext/constants_folder.py:39-115) and
 * the transpose of LIFT/D (transpose_consts_in_einsums.py:5-84). */
int dg_plan_create(const dg_plan_desc* desc, dg_plan** plan_out);

/* Frees a plan's device memory.  NULL is accepted. */
int dg_plan_destroy(dg_plan* plan);

/* Basic facts about a plan: number of fields, Np, Nfp, K. */
int dg_plan_info(const dg_plan* plan, int32_t* nfields, int32_t* np,
                 int32_t* nfp, int64_t* nelements);

/* The FP64 kernel variant the plan runs (DG_KERNEL_MFMA or DG_KERNEL_LDS;
 * DG_KERNEL_AUTO resolves per equation and order). */
int dg_plan_variant(const dg_plan* plan, int32_t* variant);

/* One right-hand-side evaluation, rhs = L(q): the call `compiled_rhs(*args)`
 * (run/__main__.py:76,109,126) of an array context, i.e. wave_operator
 * (equations/wave.py:121-200), EulerOperator.operator (equations/euler.py:
 * 176-179) or MaxwellOperator.operator (equations/maxwell.py:69-70).
 * The kernels gather neighbour traces of fields_in while they write
 * fields_out, so an output field that shares memory with an input field or
 * with another output field returns DG_ERR_INVALID_ARGUMENT (also for
 * dg_rhs_mixed). */
int dg_rhs(const dg_plan* plan, int32_t precision,
           const double* const* fields_in, double* const* fields_out,
           void* stream);

/* Mixed-precision right-hand side (precision DG_PREC_FP32 or DG_PREC_BF16X3):
 * the same operator as dg_rhs on FP32 state, FP32 RHS out.  The D and LIFT
 * contractions run on the matrix cores in FP32 (or a bf16x3 split) with FP32
 * accumulation; geometric factors, the numerical flux (Euler pressure
 * included) and the combination of reference into physical derivatives run
 * in FP64.  Triangle plans (dim = 2) take FP32 state and RHS under both
 * precisions and compute everything in FP64.  The kernels that run the
 * direct flux (FP32 at every order, BF16X3 at P3) return
 * DG_ERR_UNSUPPORTED where an FP32 field reaches 4 GiB (32-bit buffer
 * offsets); the wave ones LIFT the velocity flux in reference normals and
 * also refuse a plan whose face geometry is not the affine image of the
 * reference element's (Fscale n != rx^T c_f).  The same holds for
 * dg_rhs_lserk_stage_mixed.  It has no counterpart in the reference, whose
 * array contexts evaluate the operator in the state's own precision. */
int dg_rhs_mixed(const dg_plan* plan, int32_t precision,
                 const float* const* fields_in, float* const* fields_out,
                 void* stream);

/* One stage of the five-stage, fourth-order 2N-storage Runge-Kutta scheme
 * (Carpenter & Kennedy 1994), as in libparanumal's LSERK4 that the reference
 * compares against (assets/h200nvl.patch:42):
 *   res = a*res + dt*rhs;  q = q + b*res   for every field
 * (a == 0, the first stage, does not read res).
 * The reference has no time stepper of its own. */
int dg_lserk_stage(const dg_plan* plan, double a, double b, double dt,
                   double* const* q, double* const* res,
                   const double* const* rhs, void* stream);

/* dg_rhs followed by dg_lserk_stage in one kernel, with the stage applied to
 * the RHS in the kernel's output phase (no round trip of the RHS through
 * HBM):
 *   rhs = L(q_in);  res = a*res + dt*rhs;  q_out = q_in + b*res
 * (a == 0 does not read res).  q_out and res must share no memory with
 * q_in, each other or themselves across fields (DG_ERR_INVALID_ARGUMENT
 * otherwise), since the kernel gathers neighbour traces of q_in while it
 * writes; a stepper alternates two state buffers.  Tetrahedral plans on the MFMA
 * kernel only; others return DG_ERR_UNSUPPORTED.  No counterpart in the
 * reference (see dg_lserk_stage). */
int dg_rhs_lserk_stage(const dg_plan* plan, double a, double b, double dt,
                       const double* const* q_in, double* const* q_out,
                       double* const* res, void* stream);

/* The reverse LSERK4 stage of a linear operator (wave, Maxwell) in the
 * variables of dg_adjoint_space, where the transposed RHS is the RHS:
 *   w = L(s_in);  q_hat = q_hat + dt*w;  s_out = a*s_in + b*q_hat
 * in one kernel (dg_rhs then dg_lserk_stage_adjoint, with the update applied
 * in the RHS kernel's output phase).  q_hat is updated in place; s_out and
 * q_hat must share no memory with s_in or each other
 * (DG_ERR_INVALID_ARGUMENT otherwise).  Tetrahedral plans on the MFMA kernel
 * only; others, and Euler plans, return DG_ERR_UNSUPPORTED.  No counterpart
 * in the reference (see dg_rhs_vjp). */
int dg_rhs_lserk_stage_adjoint(const dg_plan* plan, double a, double b, double dt,
                               const double* const* s_in, double* const* s_out,
                               double* const* q_hat, void* stream);

/* LSERK4 stage for the mixed form: FP64 q and res, FP32 rhs in, and the FP32
 * shadow q32 = (float)q written for the next mixed RHS evaluation.  No
 * counterpart in the reference (see dg_lserk_stage). */
int dg_lserk_stage_mixed(const dg_plan* plan, double a, double b, double dt,
                         double* const* q, double* const* res,
                         const float* const* rhs, float* const* q32,
                         void* stream);

/* dg_rhs_mixed followed by dg_lserk_stage_mixed in one kernel, the stage
 * applied to the FP32 RHS tile in the kernel's output phase:
 *   rhs = L(q32_in) (FP32 or BF16X3);  res = a*res + dt*rhs;
 *   q = q + b*res;  q32_out = (float)q
 * (a == 0 does not read res).  The FP64 q and res are updated in place; the
 * kernel gathers neighbour traces of q32_in while it writes, so q32_out, q
 * and res must share no memory with q32_in or with each other
 * (DG_ERR_INVALID_ARGUMENT otherwise), and a stepper alternates two FP32
 * shadow buffers.  Tetrahedral plans only; triangle plans return
 * DG_ERR_UNSUPPORTED.  No counterpart in the reference (see
 * dg_lserk_stage). */
int dg_rhs_lserk_stage_mixed(const dg_plan* plan, int32_t precision, double a,
                             double b, double dt, const float* const* q32_in,
                             float* const* q32_out, double* const* q,
                             double* const* res, void* stream);

/* Vector-Jacobian product of the FP64 right-hand side at state q:
 *   q_bar = (dL/dq)^T rhs_bar
 * the exact transpose of dg_rhs for the linear wave and Maxwell operators
 * (q may then be NULL) and of dg_rhs linearised at q for Euler.  `workspace`
 * is nfields device buffers of (K, Np) float64 (overwritten).  Wave and
 * Maxwell on tetrahedra use the identity L^T = R M L R M^{-1} of their upwind
 * weak forms (see dg_adjoint_space): R M^{-1} rhs_bar into the workspace,
 * then the forward kernel with R M applied to its output (fused into the
 * MFMA kernel; a separate pass after the LDS-staged one).  Euler and
 * triangle plans: M_ref^{-1} rhs_bar, then an adjoint kernel (Euler on
 * tetrahedra: the MFMA adjoint kernel on 16-element tiles).  q_bar may
 * be rhs_bar, but not the workspace or q: the second kernel gathers
 * neighbour values from those while it writes q_bar.  The reference has
 * no adjoint: its jax:jit array context (run/__main__.py:213) is where
 * jax.vjp of the same RHS would come from.  Needs a plan created with
 * mass_inv and face_mass, else DG_ERR_UNSUPPORTED. */
int dg_rhs_vjp(const dg_plan* plan, const double* const* q,
               const double* const* rhs_bar, double* const* q_bar,
               double* const* workspace, void* stream);

/* dg_rhs_vjp with float32 state, cotangent, result and workspace (the FP32
 * adjoint); the arithmetic is FP64 as in dg_rhs_vjp.  No counterpart in the
 * reference (see dg_rhs_vjp). */
int dg_rhs_vjp_fp32(const dg_plan* plan, const float* const* q,
                    const float* const* rhs_bar, float* const* q_bar,
                    float* const* workspace, void* stream);

/* dg_rhs_vjp without a workspace argument and without allocating: no
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a static device array
 * of 128 MiB, one per GPU the library runs on, held until the process ends;
 * plan creation looks up its address on the plan's device.  The call
 * launches the same kernels as dg_rhs_vjp and calls no hipMalloc, hipFree or
 * other allocating routine.  Limit: nfields x K x Np doubles (each field
 * rounded up to 256 bytes) must fit in 128 MiB, else DG_ERR_INVALID_ARGUMENT
 * (for example wave P3 on 162,000 tets fits; Maxwell P4 on 93,750 tets does
 * not: use dg_rhs_vjp there).  Every call on one GPU shares that GPU's
 * array, so calls that overlap in time on one GPU (other streams, other
 * threads) race; the array keeps the last call's contents, and nothing in it
 * needs to start at zero.  Aliasing rules and errors as dg_rhs_vjp. */
int dg_rhs_vjp_noalloc(const dg_plan* plan, const double* const* q,
                       const double* const* rhs_bar, double* const* q_bar,
                       void* stream);

/* Reverse sweep of one LSERK4 stage (the transpose of dg_lserk_stage):
 *   q_bar = q_bar + dt*w;   s = a*s + b*q_bar
 * where w = dg_rhs_vjp(q_k, s) of the stage.  w may be NULL (the dt*w term is
 * then skipped) and a == 0 does not read s, so the sweep of a step starts with
 * (a=0, b=b_5, w=NULL) and runs stages k = 5..1 with (a_k, b_{k-1}, w_k).
 * No counterpart in the reference (see dg_lserk_stage). */
int dg_lserk_stage_adjoint(const dg_plan* plan, double a, double b, double dt,
                           double* const* q_bar, double* const* s,
                           const double* const* w, void* stream);

/* dg_lserk_stage_adjoint on float32 q_bar, s and w.  No counterpart in the
 * reference (see dg_lserk_stage). */
int dg_lserk_stage_adjoint_fp32(const dg_plan* plan, double a, double b, double dt,
                                float* const* q_bar, float* const* s,
                                const float* const* w, void* stream);

/* Change of variables of the linear operators' adjoint (wave, Maxwell; FP64):
 *   to_adjoint != 0:  out = R M^{-1} in;     to_adjoint == 0:  out = R M in
 * with M the global mass matrix (|J| M_ref per element) and R the sign flip
 * of the velocity (wave) or magnetic (Maxwell) fields; in and out may be the
 * same buffers.  In these variables the transpose of the RHS is the RHS
 * itself -- L^T = R M L R M^{-1}, see dg_rhs_vjp -- so a reverse LSERK4 sweep
 * runs on dg_rhs and dg_lserk_stage_adjoint between one transform in and one
 * out.  Needs a plan created with mass_inv and face_mass; Euler plans return
 * DG_ERR_UNSUPPORTED.  No counterpart in the reference (see dg_rhs_vjp). */
int dg_adjoint_space(const dg_plan* plan, int32_t to_adjoint, const double* const* in,
                     double* const* out, void* stream);

/* dg_adjoint_space on float32 fields (FP64 arithmetic).  An FP32 reverse
 * sweep runs in these variables on dg_rhs_mixed (DG_PREC_FP32) and
 * dg_lserk_stage_adjoint_fp32.  No counterpart in the reference (see
 * dg_rhs_vjp). */
int dg_adjoint_space_fp32(const dg_plan* plan, int32_t to_adjoint, const float* const* in,
                          float* const* out, void* stream);

/* Message for the last failing call on this thread ("" if none). */
const char* dg_last_error_string(void);

#ifdef __cplusplus
}
#endif

#endif /* DGFEM_H */
