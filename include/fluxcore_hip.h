/* fluxcore_hip.h — C ABI of libfluxcore_hip.so (MI355X / gfx950).
 *
 * Conventions
 *  - All pointers are device pointers (hipMalloc / torch CUDA tensors) unless
 *    stated otherwise; sizes are element counts.
 *  - Populations are SoA f[q*N + flat], flat = (i*ny + j)*nz + k, the layout of
 *    the reference's f[19, nx, ny, nz] tensor (solver/lbm_kernels.cu:6-9).
 *  - Every int-returning entry point returns 0 on success; on failure it
 *    returns non-zero and fc_last_error() holds the message (per thread).
 *  - Work is enqueued on `stream` and returns without synchronising.
 *  - No entry point keeps a reference to its arguments; scratch memory is
 *    passed in by the caller (workspace arguments).
 *
 * Each declaration names the reference interface it replaces (path:line in
 * the fluxcore3D tree).
 */
#ifndef FLUXCORE_HIP_H
#define FLUXCORE_HIP_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *fc_last_error(void);
int fc_version(void);

/* ---- D3Q19 kernels, one per pybind function of lbm_kernels.cu:1347-1386 ---- */

/* macro_collide_bgk (lbm_kernels.cu:628-647): f -> f_post, pre-collision rho,u,v,w;
 * solid nodes get rest weights and (1,0,0,0). */
int fc_lbm_macro_collide_bgk(const float *f, float *f_post, float *rho, float *u, float *v, float *w,
                             const uint8_t *solid, float omega, int64_t N, hipStream_t stream);
/* macro_collide_trt (lbm_kernels.cu:649-669) */
int fc_lbm_macro_collide_trt(const float *f, float *f_post, float *rho, float *u, float *v, float *w,
                             const uint8_t *solid, float omega_plus, float omega_minus, int64_t N,
                             hipStream_t stream);
/* macro_collide_mrt (lbm_kernels.cu:671-692); S_rates: 19 device floats by moment row */
int fc_lbm_macro_collide_mrt(const float *f, float *f_post, float *rho, float *u, float *v, float *w,
                             const uint8_t *solid, const float *S_rates, int64_t N, hipStream_t stream);
/* macro_collide_mrt_smag (lbm_kernels.cu:1120-1145) */
int fc_lbm_macro_collide_mrt_smag(const float *f, float *f_post, float *rho, float *u, float *v, float *w,
                                  const uint8_t *solid, const float *S_rates, float Cs_dx2, float omega0,
                                  int64_t N, hipStream_t stream);
/* stream_pull (lbm_kernels.cu:695-708): pull with periodic wrap / copy-self */
int fc_lbm_stream_pull(const float *f_in, float *f_out, int nx, int ny, int nz, int periodic_x, int periodic_y,
                       int periodic_z, hipStream_t stream);
/* sponge (lbm_kernels.cu:711-727) */
int fc_lbm_sponge(float *f, const float *rho, const float *u, const float *v, const float *w, const float *sigma,
                  int64_t N, hipStream_t stream);
/* ibb (lbm_kernels.cu:730-754): Bouzidi per link, writes f_stream[dir_i][fluid_flat] */
int fc_lbm_ibb(const float *f_post, float *f_stream, const int32_t *fluid_flat, const int32_t *dir_i,
               const int32_t *dir_opp, const float *w1, const float *w2, const uint8_t *q_ge_half,
               const int32_t *ff_flat, const uint8_t *ff_valid, int64_t n_links, int64_t N, hipStream_t stream);
/* reset_solid (lbm_kernels.cu:757-764) */
int fc_lbm_reset_solid(float *f, const uint8_t *solid, int64_t N, hipStream_t stream);
/* zou_he_bc (lbm_kernels.cu:975-996): regularised Zou/He on one plane */
int fc_lbm_zou_he_bc(float *f, const uint8_t *solid, int nx, int ny, int nz, int axis, int plane_idx,
                     int interior_idx, int bc_type, float u_in_lat, float rho_out_lat, int sign,
                     hipStream_t stream);
/* convective_bc (lbm_kernels.cu:1299-1344): zero-gradient copy + mass correction.
 * The reference keeps two static device scalars; here the caller passes a
 * workspace of fc_lbm_convective_bc_workspace_bytes(n1, n2) bytes. */
size_t fc_lbm_convective_bc_workspace_bytes(int n1, int n2);
int fc_lbm_convective_bc(float *f, const uint8_t *solid, int nx, int ny, int nz, int axis, int outlet_idx,
                         int interior_idx, int unknown_cax, float rho_out_lat, int n1, int n2, void *workspace,
                         size_t workspace_bytes, hipStream_t stream);
/* convective_bc without a workspace and without any allocation. No counterpart in
 * the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * partial sums live in a 64 KiB static device array of the library (one per GPU,
 * held until the process ends), so the outlet plane is limited to n1 * n2 <=
 * 1,048,576 nodes (larger planes return an error). Calls that overlap on one GPU
 * (different streams) race on that array. Same results as fc_lbm_convective_bc. */
int fc_lbm_convective_bc_noalloc(float *f, const uint8_t *solid, int nx, int ny, int nz, int axis, int outlet_idx,
                                 int interior_idx, int unknown_cax, float rho_out_lat, int n1, int n2,
                                 hipStream_t stream);

/* ---- fused in-place step (AA pattern): replaces the kernel sequence of
 *      LBMCUDAUpgrade._step_cuda (lbm_cuda_ext.py:445-517) ---- */
typedef struct fc_aa_tables {
  const uint8_t *nflag;     /* [N] 1 solid, 2 special, bits 3..5 first BC plane + 1 */
  const int32_t *wave_base; /* [ceil(N/64)] special nodes before each 64-node group */
  const int32_t *wave_meta; /* [ceil(n_lanes/64) + 1][4] per wavefront of 64 lanes: first link, first
                               second-node link, first publish entry, first solid-source entry (last: totals) */
  const int32_t *lk_code;   /* [n_links] lane | t << 6 | opp(t) << 11 | second node << 16
                               | override << 17 | second-node index in the wavefront << 18 */
  const int32_t *s_code;    /* [n_src] lane | q << 6 | override << 17 (input w_q after step 1) */
  const int32_t *f_code;    /* [n_publish] lane | p << 6 (post-collision f_p -> B[ff_idx]) */
  const int32_t *lane_nodes; /* [n_lanes] node of each lane, 64 lanes per wavefront: the fluid nodes
                               (dense storage), the nodes owning slots (compact storage), or NULL for
                               one lane per node */
  int64_t n_lanes;
  const int32_t *sp_mask;   /* [n_special][4] link, solid-source, second-node, publish masks */
  const int32_t *sp_base;   /* [n_special][4] first link, first second-node link, first publish entry */
  const float *w1;          /* [n_links] Bouzidi weights */
  const float *w2;          /* [n_links], or NULL: w2 = 1 - w1 (the reference's weights, exactly) */
  float *A;                 /* [n_links] own-node part of each link value */
  float *B;                 /* [2][n_v] second-node part, by step parity */
  const int32_t *ff_idx;    /* [n_publish] slot in B written by each publishing node, or NULL: the slot
                               minus the group's first second-node slot is f_code >> 11 (signed) */
  int64_t n_v;
  /* Compact storage (NULL: dense, every padded node owns slots). Per padded row
   * R = I * PY + J and word w = K / 32: {mask, start}, bit b of mask set when node
   * K = 32 w + b owns slots, start = nodes owning slots before the word (padded C
   * order); storage index = start + popcount(mask & ((1 << b) - 1)). Nodes without
   * slots stream w_q. */
  const int32_t *row_meta;  /* [PX * PY * wpr][2] */
  int wpr;                  /* words per row, (PZ + 31) / 32 */
  int64_t n_store;          /* stride of the 19 population arrays ([19][n_store]): at least the number of
                               nodes owning slots, at most the dense slot count */
} fc_aa_tables;

typedef struct fc_aa_collision {
  int kind;            /* 0 BGK, 1 TRT, 2 MRT, 3 MRT + Smagorinsky */
  float om1, om2;      /* BGK omega | TRT omega+, omega- | Smagorinsky omega0 */
  float S[19];         /* MRT rates by moment row */
  float cs_dx2;        /* (Cs dx)^2 */
  float force[3];      /* Guo body force (lattice units) */
  int use_force;
  const float *sponge; /* device sigma profile along sponge_axis, or NULL */
  int sponge_axis;
} fc_aa_collision;

/* storage: 0 FP32 populations, 1 FP16 deviations f - w_i, 2 BF16 deviations.
 * px/py/pz: periodic axes (non-periodic axes carry a one-cell ghost layer). */
int64_t fc_aa_slot_count(int nx, int ny, int nz, int px, int py, int pz);
/* One fused step (replaces the sequence of lbm_cuda_ext.py:445-517). rho/u/v/w: pre-collision
 * macros (all four or none); solid_macros = 0 skips the constant solid-node values (1, 0, 0, 0)
 * when the caller passes the same buffers as on the previous step. */
int fc_aa_step(void *pop, int storage, int nx, int ny, int nz, int px, int py, int pz, const fc_aa_tables *tables,
               const fc_aa_collision *collision, int parity, int first, float *rho, float *u, float *v, float *w,
               int solid_macros, hipStream_t stream);
int fc_aa_materialize(const void *pop, int storage, int nx, int ny, int nz, int px, int py, int pz,
                      const fc_aa_tables *tables, int parity, int first, float *f_out, hipStream_t stream);
int fc_aa_load(void *pop, int storage, int nx, int ny, int nz, int px, int py, int pz, const fc_aa_tables *tables,
               const float *f_in, hipStream_t stream);
/* inlet/outlet planes on the AA state (zou_he_bc / convective_bc semantics) */
int fc_aa_zou_he(void *pop, int storage, int nx, int ny, int nz, int px, int py, int pz, const fc_aa_tables *tables,
                 int parity, int first, int bc_order, int axis, int plane_idx, int interior_idx, int bc_type,
                 float u_in_lat, float rho_out_lat, int sign, hipStream_t stream);
/* device workspace of fc_aa_convective for an n1 x n2 outlet plane: per-block partial
 * sums and a plane buffer of the outlet inputs and their slot indices (about 152 bytes
 * per plane node) */
size_t fc_aa_convective_workspace_bytes(int n1, int n2);
int fc_aa_convective(void *pop, int storage, int nx, int ny, int nz, int px, int py, int pz,
                     const fc_aa_tables *tables, int parity, int first, int bc_order, int axis, int outlet_idx,
                     int interior_idx, int unknown_cax, float rho_out_lat, void *workspace, size_t workspace_bytes,
                     hipStream_t stream);

/* ---- Bouzidi link set and wall distances (pre_post/sdf_tracing.py:20-259) ---- */
/* build_links_from_mask_gpu (sdf_tracing.py:20-85): per node, bit q of mask when
 * x + c_q is solid, bit q of ffok when x - c_q is a valid fluid node. */
int fc_links_mask(const uint8_t *solid, int nx, int ny, int nz, int px, int py, int pz, uint32_t *mask,
                  uint32_t *ffok, hipStream_t stream);
/* bfl_q_from_sdf + bfl_weights_from_q (sdf_tracing.py:127-205) with x0, d, L of
 * build_ibb_from_sdf_gpu (:208-247); px/py/pz = 1 wraps the SDF samples on that
 * axis (the reference clamps; 0 reproduces it) */
int fc_links_q(const float *phi, int nx, int ny, int nz, float ox, float oy, float oz, float pitch, int px, int py,
               int pz, const int64_t *fluid_flat, const int64_t *dir_i, int64_t n, int n_coarse, int n_bisect,
               float *q, float *w1, float *w2, uint8_t *q_ge_half, hipStream_t stream);
/* sdf_trilinear (sdf_tracing.py:88-125): phi at points x[n][3] (world units,
 * origin (ox, oy, oz), spacing pitch), grid coordinates clamped to [0, n - 1.001] */
int fc_sdf_trilinear(const float *phi, int nx, int ny, int nz, float ox, float oy, float oz, float pitch,
                     const float *x, int64_t n, float *out, hipStream_t stream);
/* bfl_q_from_sdf (sdf_tracing.py:127-184) for arbitrary rays: start points
 * x0[n][3], unit directions d[n][3], lengths L[n] -> q[n] */
int fc_bfl_q_rays(const float *phi, int nx, int ny, int nz, float ox, float oy, float oz, float pitch,
                  const float *x0, const float *d, const float *L, int64_t n, int n_coarse, int n_bisect, float *q,
                  hipStream_t stream);
/* bfl_weights_from_q (sdf_tracing.py:187-205): q >= 1/2, w1, w2 */
int fc_bfl_weights(const float *q, int64_t n, uint8_t *q_ge_half, float *w1, float *w2, hipStream_t stream);


/* ---- geometric multigrid (mg_kernels.cu / nlevel_mg_cuda.py) ---- */
/* mop_fused (mg_kernels.cu:133-174): Lo = T + dt L(T) with the reference ghosts */
int fc_mg_mop_fused(const float *T, const float *aE, const float *aW, const float *aN, const float *aS,
                    const float *aU, const float *aD, const float *diagL, float *Lo, double dt, int nx, int ny, int nz,
                    int ax, int out_idx, hipStream_t stream);
/* jacobi_smooth (mg_kernels.cu:177-197): x += inv_diagM (b - Lo) */
int fc_mg_jacobi_smooth(float *x, const float *b, const float *Lo, const float *inv_diagM, int64_t n,
                        hipStream_t stream);
/* _build_coeffs_full (nlevel_mg_cuda.py:418-507): aE..aD, diagL of one level */
int fc_mg_build_coeffs(const float *k, const float *rho_cp, const float *u, const float *v, const float *w, int nx,
                       int ny, int nz, int ax, int out_idx, double dx, double s, float *aE, float *aW, float *aN,
                       float *aS, float *aU, float *aD, float *diagL, hipStream_t stream);
/* _update_inv_diag (nlevel_mg_cuda.py:262-267) */
int fc_mg_inv_diag(const float *diagL, const float *mask, double dt, double omega, float *inv_diagM, int64_t n,
                   hipStream_t stream);
/* _restrict_2x (nlevel_mg_cuda.py:512-519); threshold=1 gives the coarse mask (>= 0.999);
 * threshold=2 (no counterpart in the reference) marks a coarse cell as unknown where
 * any child is, for systems whose other cells are decoupled by zero conductance */
int fc_mg_restrict_field(const float *src, float *dst, int nx, int ny, int nz, int threshold, hipStream_t stream);

/* One level's coefficients. fx, fy, fz (all set or all null) hold the face form
 * written by fc_mg_build_faces: per axis, the coefficient of the face between a
 * cell and its +1 neighbour, the last plane holding the low-boundary coefficient
 * where the low ghost is the cell itself. When set, the kernels below read them
 * instead of aE..aD and diagL (3 arrays instead of 7 per cell). */
typedef struct fc_mg_coef {
  const float *aE, *aW, *aN, *aS, *aU, *aD, *diagL, *invD;
  const float *fx, *fy, *fz;
  /* NULL, or uint8 [ceil(nx ny nz / 64)]: entry s covers the 64 cells of flat index
   * 64 s .. 64 s + 63; 0 where those cells all have zero coefficients and their
   * right-hand side and iterate are 0. The sweeps, the residual restriction,
   * fc_mg_apply_A64 and fc_mg_true_residual skip such runs and leave their outputs
   * there as they are (zeros), which gives the same bits as visiting them. */
  const uint8_t *rows;
} fc_mg_coef;

/* No reference counterpart (a storage form of _build_coeffs_full's output,
 * nlevel_mg_cuda.py:418-507): writes the face form of coef's per-cell arrays and
 * adds to *mismatches the number of cells whose six coefficients or diagL, read
 * back through it, differ in any bit (zero for conduction with uniform rho*c_p and
 * no advection). Use the face form only when that count is zero. */
int fc_mg_build_faces(const fc_mg_coef *coef, int nx, int ny, int nz, int ax, int out_idx, float *fx, float *fy,
                      float *fz, unsigned *mismatches, hipStream_t stream);

/* fused _Mop + _smooth_step (nlevel_mg_cuda.py:311-372): x_out = x + invD (b - A x) */
int fc_mg_sweep(const float *x, float *x_out, const float *b, const fc_mg_coef *coef, double dt, int nx, int ny,
                int nz, int ax, int out_idx, int zero_guess, hipStream_t stream);
/* the first two _smooth_step sweeps of _vcycle from x = 0 (nlevel_mg_cuda.py:285-287) in one
 * pass: x_out = x1 + invD (b - A x1), x1 = invD b formed at the stencil points; the same
 * bits as fc_mg_sweep with zero_guess=1 followed by fc_mg_sweep */
int fc_mg_sweep2_from_zero(float *x_out, const float *b, const fc_mg_coef *coef, double dt, int nx, int ny, int nz,
                           int ax, int out_idx, hipStream_t stream);
/* residual + _restrict_into of _vcycle (nlevel_mg_cuda.py:289-295) */
int fc_mg_residual_restrict(const float *x, const float *b, const float *mask, const fc_mg_coef *coef, double dt,
                            int nx, int ny, int nz, int ax, int out_idx, float *b_coarse, hipStream_t stream);
/* _prolong_2x + correction of _vcycle (nlevel_mg_cuda.py:300-303,531-537); rows (NULL: none)
 * is the fine level's table of skipped 64-cell runs, as in fc_mg_coef */
int fc_mg_prolong_correct(const float *x_coarse, float *x, const float *mask, int nx, int ny, int nz, int cx, int cy,
                          int cz, float alpha, const uint8_t *rows, hipStream_t stream);
/* coarsest-level sweeps of _vcycle (nlevel_mg_cuda.py:279-283) in one workgroup */
int fc_mg_coarse_solve(float *x, float *tmp, const float *b, const fc_mg_coef *coef, double dt, int nx, int ny,
                       int nz, int ax, int out_idx, int sweeps, int zero_guess, hipStream_t stream);
/* residual norm of solve() (nlevel_mg_cuda.py:133-135), FP64 accumulation */
int fc_mg_norm2_f32(const float *x, int64_t n, double *partial, double *scal, int slot, hipStream_t stream);
int fc_mg_residual_f32(const float *b, const float *Lo, const float *mask, float *r, int64_t n, double *partial,
                       double *scal, int slot, hipStream_t stream);

/* FP64 flexible-CG outer loop replacing pcg_mg (mg_pcg.py:33-97); partial holds
 * 3 * fc_mg_partials_count(n) doubles, scal fc_mg_scalar_slots() doubles */
int fc_mg_partials_count(int64_t n);
int fc_mg_scalar_slots(void);
/* chebyshev_smooth (mg_pcg.py:195-247) building blocks on one level:
 * lambda_max = max(diagL dt + 1) into out[0] (partial: fc_mg_cheb_partials_count(n) floats);
 * first step r = (b - Lo) mask, p = invD r, x_new = x + p;
 * later steps p = rho_new (s invD r + rho_prev p), x_new += p, and x_final = x_new mask
 * when x_final is not NULL (FP32, the scalars rounded to FP32 as the reference's tensor ops do) */
int fc_mg_cheb_partials_count(int64_t n);
int fc_mg_cheb_lambda_max(const float *diagL, int64_t n, double dt, float *partial, float *out, hipStream_t stream);
int fc_mg_cheb_first(const float *x, const float *b, const float *Lo, const float *invD, const float *mask, float *p,
                     float *x_new, int64_t n, hipStream_t stream);
int fc_mg_cheb_step(float *x_new, const float *b, const float *Lo, const float *invD, const float *mask, float *p,
                    double rho_new, double s, double rho_prev, float *x_final, int64_t n, hipStream_t stream);
int fc_mg_apply_A64(const double *p, double *Ap, const double *r, const float *mask, const fc_mg_coef *coef,
                    double dt, int nx, int ny, int nz, int ax, int out_idx, double *partial, double *scal,
                    hipStream_t stream);
/* rows (NULL: none): one entry per run of row_len consecutive cells, ceil(n / row_len)
 * entries (the finest level's fc_mg_coef rows with row_len = 64); the cells of runs
 * whose entry is 0 are left as they are */
int fc_mg_update_xr(double *x, double *r, float *r32, const double *p, const double *Ap, int64_t n,
                    const uint8_t *rows, int row_len, double *partial, double *scal, hipStream_t stream);
int fc_mg_new_direction(double *p, const float *z, const double *Ap_old, int64_t n, int first, const uint8_t *rows,
                        int row_len, double *partial, double *scal, hipStream_t stream);
int fc_mg_true_residual(const double *x, const double *b, double *r, float *r32, const float *mask,
                        const fc_mg_coef *coef, double dt, int nx, int ny, int nz, int ax, int out_idx,
                        double *partial, double *scal, hipStream_t stream);
/* Stopping test on the device, in place of the host test after every iteration of
 * pcg_mg (mg_pcg.py:69-72): scal[stop_slot] = 1 once sqrt(scal[rr_slot]) / bnorm < tol
 * or is not finite, sticky until a call with clear != 0. With the flag set,
 * fc_mg_new_direction and fc_mg_update_xr (stop slot fc_mg_stop_slot()) and
 * fc_gcr_step (fc_gcr_stop_slot()) leave x, r and the directions unchanged, so the
 * host may queue iteration k+1 before it reads iteration k's result. */
int fc_solver_stop_check(double *scal, int rr_slot, int stop_slot, double bnorm, double tol, int clear,
                         hipStream_t stream);
int fc_mg_stop_slot(void);

/* ---- thermal step kernels (thermal_kernels.cu:198-263) ---- */
int fc_thermal_rhs(const float *T_old, const float *T0, const float *src, const float *LT0, const float *mask,
                   float *rhs, float *theta0, double dt, int64_t n, hipStream_t stream);
int fc_thermal_reconstruct(const float *T0, const float *theta, float *T_new, int nx, int ny, int nz, int ax,
                           int out_idx, hipStream_t stream);
size_t fc_thermal_metrics_workspace_bytes(int64_t n);
int fc_thermal_convergence_metrics(const float *T_new, const float *T_old, const uint8_t *solid, int64_t n,
                                   float *out_solid, float *out_fluid, void *workspace, size_t workspace_bytes,
                                   hipStream_t stream);
size_t fc_thermal_check_workspace_bytes(int64_t n);
int fc_thermal_physics_check(const float *T, double T_min, double T_max, int64_t n, int *flag, void *workspace,
                             size_t workspace_bytes, hipStream_t stream);

/* ---- learned preconditioner (csrc/conv.hip, csrc/unet.hip) ----
 * No reference counterpart: the reference's thermal solve ends at the
 * multigrid-preconditioned Krylov methods (solver/mg_pcg.py:33-187). These
 * kernels run the U-Net preconditioner used inside flexible CG.
 * Activations are channels-last BF16 [B, D, H, W, C] (C a multiple of 8);
 * weights BF16 [cout][fc_conv_kpad(cin)], k = tap * cin + ci with
 * tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padded; FP32 accumulation. */
int fc_conv_kpad(int cin);
/* y = act(conv3x3x3(concat(up0 ? nearest_up2(x0) : x0, x1))), zero padding;
 * act 1 = leaky ReLU(slope); output BF16 (y_bf16) and/or FP32 (y_f32). */
int fc_conv3d_fwd(const void *x0, int c0, int up0, const void *x1, int c1, const void *w, int cout, int B, int D,
                  int H, int W, int act, float slope, void *y_bf16, float *y_f32, hipStream_t stream);
/* weights of the data gradient: w_flip[ci][t*cout + co] = w[co][(26 - t)*cin + ci] */
int fc_conv3d_flip_weights(const void *w, int cout, int cin, void *w_flip, hipStream_t stream);
size_t fc_conv3d_dw_workspace_bytes(int cout, int cin, int n_chunks);
/* dw[co][t][ci] (FP32) = sum over positions of dy[p, co] * x[p + o_t, ci] */
int fc_conv3d_bwd_weight(const void *x0, int c0, int up0, const void *x1, int c1, const void *dy, int cout, int B,
                         int D, int H, int W, int n_chunks, float *dw, void *workspace, size_t ws_bytes,
                         hipStream_t stream);
/* x[p] = {r[p] sqrt(n / scal[slot]), feat1[p], mask[p], 0 x 5} (BF16, 8 channels; residual at unit RMS) */
int fc_unet_pack_input(const float *r, const float *feat1, const float *mask, const double *scal, int slot,
                       int64_t n, void *out, hipStream_t stream);
int fc_unet_avgpool2(const void *in, int B, int D, int H, int W, int C, void *out, hipStream_t stream);
/* z[p] = sqrt(scal[slot] / n) * y[p * cy] * mask[p] * d[p] (mask, d optional) */
int fc_unet_unpack(const float *y, int cy, const float *mask, const float *d, const double *scal, int slot,
                   int64_t n, float *z, hipStream_t stream);
/* out[p] = r[p] * d[p] */
int fc_unet_scale(const float *r, const float *d, int64_t n, float *out, hipStream_t stream);
/* z += s y[:, 0] mask d (network correction added to a V-cycle's output) */
int fc_unet_unpack_add(const float *y, int cy, const float *mask, const float *d, const double *scal, int slot,
                       int64_t n, float *z, hipStream_t stream);
/* out = (r - Lo) mask d: the scaled residual left by z, Lo = A z */
int fc_unet_residual_scale(const float *r, const float *Lo, const float *mask, const float *d, int64_t n,
                           float *out, hipStream_t stream);

/* ---- flexible GCR(m) in FP64 (csrc/krylov.hip) ----
 * No reference counterpart: the minimal-residual Krylov method used with the
 * learned preconditioner (the reference's Krylov solvers are mg_pcg.py:33-187). */
typedef struct fc_gcr_stencil {
  const float *aE, *aW, *aN, *aS, *aU, *aD, *diagL, *mask; /* finest-level operator, as fc_mg_coef */
  double dt;
  int nx, ny, nz, ax, out_idx;
  const float *fx, *fy, *fz; /* face form (fc_mg_build_faces) or all null */
} fc_gcr_stencil;
int fc_gcr_scalar_slots(void);
int fc_gcr_stop_slot(void);
int fc_gcr_max_directions(void);
int fc_gcr_partials_count(int64_t n);
/* p = z, w = A p, orthogonalise (w, p) against the m stored pairs (W_j, P_j) whose W_j.W_j sit in
 * scal[ww_slot[j]], alpha = r.w / w.w, x += alpha p, r -= alpha w, r32 = (float) r;
 * scal[2] = r.r, scal[store_slot] = w.w (slots 16..23 hold the W.W history). */
int fc_gcr_step(const float *z32, const fc_gcr_stencil *stencil, double *p, double *w, const double *const *W,
                double *const *P, const int *ww_slot, int m, int store_slot, double *x, double *r, float *r32,
                double *partial, double *scal, hipStream_t stream);

/* ---- restarted flexible GMRES(m) in FP64 (csrc/krylov.hip) ----
 * No reference counterpart: a minimal-residual Krylov method for the conduction system with
 * the V-cycle (or the learned preconditioner) as a variable preconditioner; the reference's
 * Krylov solvers are mg_pcg.py:33-187. The basis V_0..V_m stays unnormalised (FP64), the
 * preconditioned vectors Z_j are the preconditioner's FP32 output. scal holds
 * fc_fgmres_scalar_slots() doubles: w.V_i at fc_fgmres_dot_slot(i), ||V_i||^2 at
 * fc_fgmres_norm_slot(i); partial holds fc_fgmres_partials_count(n) doubles. */
int fc_fgmres_max_basis(void);
int fc_fgmres_scalar_slots(void);
int fc_fgmres_dot_slot(int i);
int fc_fgmres_norm_slot(int i);
int fc_fgmres_partials_count(int64_t n);
/* one Arnoldi step: V[j+1] = A z32 (with w.V_i, i <= j), then V[j+1] -= sum_i (w.V_i / V_i.V_i) V_i,
 * v32 = (float) V[j+1], ||V[j+1]||^2; needs ||V_i||^2, i <= j, in scal */
int fc_fgmres_arnoldi(const float *z32, const fc_gcr_stencil *stencil, double *const *V, int j, float *v32,
                      double *partial, double *scal, hipStream_t stream);
/* x += sum_{i<k} a[i] Z[i]; a is a host array of k doubles */
int fc_fgmres_update_x(double *x, const float *const *Z, const double *a, int k, int64_t n, hipStream_t stream);
/* the Givens rotations of step j on the device (one thread), so the host need not wait for
 * each step: state holds fc_fgmres_givens_slots() doubles (R column-major with leading
 * dimension *ld from offset *r, g from *g, the estimate |g_j+1| / bnorm at *est) */
int fc_fgmres_givens_slots(void);
int fc_fgmres_givens_layout(int *r, int *g, int *est, int *ld);
int fc_fgmres_givens(const double *scal, double *state, int j, double bnorm, hipStream_t stream);

/* ---- MG-preconditioned BiCGSTAB (bicgstab_mg, mg_pcg.py:103-175; its branch of patch_mg_pcg :250-275) ----
 * FP64 vectors, FP32 preconditioned vectors; the scalars alpha, beta, omega are formed by the
 * host from the dot products, in the reference's order. */
/* out = (z + dt L z) [* mask if use_mask], FP64 from FP32 z (the reference's _Mop keeps masked rows) */
int fc_krylov_apply(const float *z32, const fc_gcr_stencil *stencil, double *out, int use_mask, hipStream_t stream);
/* scal[slot0 + i] = a[i] . b[i], i < npairs (1..3); partial: fc_gcr_partials_count(n) doubles */
int fc_krylov_dots(const double *const *a, const double *const *b, int npairs, int64_t n, double *partial,
                   double *scal, int slot0, hipStream_t stream);
/* p = (beta p + r) - beta omega v; p32 = (float) (p * mask) */
int fc_bicgstab_direction(double *p, const double *r, const double *v, const float *mask, double beta, double omega,
                          float *p32, int64_t n, hipStream_t stream);
/* s = (r - alpha v) * mask; s32 = (float) s */
int fc_bicgstab_s(double *s, const double *r, const double *v, const float *mask, double alpha, float *s32, int64_t n,
                  hipStream_t stream);
/* x = ((x + alpha p_hat) + omega s_hat) * mask and r = (s - omega t) * mask; with s_hat = NULL
 * only x = (x + alpha p_hat) * mask */
int fc_bicgstab_update(double *x, double *r, const double *s, const double *t, const float *p_hat,
                       const float *s_hat, const float *mask, double alpha, double omega, int64_t n,
                       hipStream_t stream);

/* ---- D3Q7 fluid temperature and the solid conduction step (csrc/d3q7.hip) ----
 * No counterpart in the reference: its temperature is an implicit finite-volume
 * upwind operator solved after the flow is frozen (nlevel_mg_cuda.py:418-507,
 * LBM_CHT.py:2208-2359, cht_worker.py:23-165). These carry the fluid temperature
 * every flow step and hand the energy crossing the fluid-solid interface to the
 * solid, whose conduction step is solved with the multigrid entry points above.
 * flag: 0 fluid, 1 conducting solid (temperature T_solid, solved apart), 2 adiabatic
 * wall, 3 conducting solid carried on the lattice (u = 0, relaxation omega_s:
 * conjugate conduction in the same launch), | 16 inlet (held at T_in).
 * g_in, g_out: [7][N] post-collision populations (0, +x, -x, +y, -y, +z, -z),
 * N = nx ny nz < 2^31 in C order; Q (may be NULL): fluid heat source per step;
 * acc (may be NULL): per-node energy the solid receives, accumulated. nbmask (may be NULL,
 * static geometry), bits q - 1 for q = 1..6 and neighbours inside the domain: on a node of
 * flag kind 1, x + c_q is a fluid node off the inlet; on every other node, x - c_q is a
 * wall or a conducting solid (flag 1 or 2). Only the neighbours it names are read. */
int fc_d3q7_step(const float *g_in, float *g_out, float *T_out, const float *u, const float *v, const float *w,
                 const uint8_t *flag, const uint8_t *nbmask, const float *T_solid, const float *Q, float *acc,
                 int nx, int ny, int nz, int px, int py, int pz, float omega, float omega_s, float T_in,
                 hipStream_t stream);
/* g = w_i T (1 + 4 c_i.u) */
int fc_d3q7_init(float *g, const float *T, const float *u, const float *v, const float *w, int64_t n,
                 hipStream_t stream);
/* right-hand side b = T_s + dt src + acc / rho_cp and start x0 = T_s (2 T_s - T_prev
 * when T_prev, the previous interval's start temperature, is given; 3 (T_s - T_prev) +
 * T_prev2 when the one before, T_prev2, is given too) of the solid's implicit step
 * (I + dt L) T = b on a box (origin ox, oy, oz >= -1; extents bx, by, bz) of the
 * nx x ny x nz grid; zero on box cells that are not conducting */
int fc_cht_solid_rhs(const float *T_solid, const float *T_prev, const float *T_prev2, const float *acc,
                     const float *src, const uint8_t *flag, int nx, int ny, int nz, int ox, int oy, int oz, int bx,
                     int by, int bz, double rho_cp, double dt, double *b, double *x0, hipStream_t stream);
/* T_s = x on the box's conducting cells; their acc restarts at zero */
int fc_cht_solid_store(const double *x, float *T_solid, float *acc, const uint8_t *flag, int nx, int ny, int nz,
                       int ox, int oy, int oz, int bx, int by, int bz, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* FLUXCORE_HIP_H */
