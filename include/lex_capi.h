/* C ABI of liblexhip.so: MI355X (gfx950) kernels for LEX's per-step dynamics.
 *
 * Conventions
 *  - Arrays are FP32 device pointers in the reference's padded C-order layout (x, y, z),
 *    z contiguous, ghost widths (3, 3, 1):
 *        scalar (nx+6, ny+6, nz+2)   u (nx+7, ny+6, nz+2)
 *        v      (nx+6, ny+7, nz+2)   w (nx+6, ny+6, nz+3)
 *    "interior" arrays are unpadded: scalars (nx, ny, nz), u-points (nx+1, ny, nz),
 *    v-points (nx, ny+1, nz), w-points (nx, ny, nz+1).
 *  - The base state is horizontally uniform (setup_lex.py:76-137); entry points take
 *    its profiles as z-columns of nz+2 values (*_col).
 *  - Every call is asynchronous on `stream` (a hipStream_t; NULL = default stream).
 *    The caller owns all memory; nothing is retained after return.
 *  - Return value: 0 = ok, <0 = argument/shape error, >0 = HIP error code.
 *    lex_last_error() returns the message of the last failure on this thread.
 *
 * Each declaration names the reference routine it replaces (file:line in
 * MetLab-HKUST/LEX).  The reference is pure JAX with no FFI of its own; the Python
 * modules in lex_amd/ keep the reference's function signatures and call these.
 */
#ifndef LEX_CAPI_H
#define LEX_CAPI_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lex_grid {
  int nx, ny, nz;      /* interior cells (namelist_n_constants.py:20-22) */
  float dx, dy, dz;    /* uniform spacings (namelist_n_constants.py:17-19) */
} lex_grid_t;

const char* lex_last_error(void);
int lex_abi_version(void);

/* ---- H1: halo padding --------------------------------------------------------- */
/* padding3_array / padding3u_array / padding3v_array (one_step_integration.py:950-985):
 * mode 0 scalar or w (ni, nj, nk) -> (ni+6, nj+6, nk+2); mode 1 u (ni=nx+1);
 * mode 2 v (nj=ny+1).  pad = 3.  padding_array (pressure_equations.py:111-120): pad = 1. */
int lex_pad(int mode, int ni, int nj, int nk, int pad, const float* in, float* out, void* stream);
/* set_w_bc (boundary_conditions.py:104-109), in place */
int lex_set_w_bc(const lex_grid_t* g, float* w, void* stream);

/* ---- A1: divergences ------------------------------------------------------------ */
/* get_divergence (advection.py:200-226) -> padded scalar array */
int lex_div_rho(const lex_grid_t* g, const float* rho0_col, const float* u, const float* v, const float* w,
                float* out, void* stream);
/* get_divergence2 (advection.py:229-236) -> padded scalar array */
int lex_div(const lex_grid_t* g, const float* u, const float* v, const float* w, float* out, void* stream);

/* ---- A2-A4: advection ------------------------------------------------------------ */
/* advection_scalar (advection.py:9-34); sfc_flux (nx, ny) or NULL -> out interior */
int lex_adv_scalar(const lex_grid_t* g, const float* rho0_col, const float* s, const float* u, const float* v,
                   const float* w, float weps, const float* flow_div, const float* sfc_flux, float* out, void* stream);
/* advection_u (advection.py:37-74); sfc_flux (nx, ny) or NULL -> (nx+1, ny, nz) */
int lex_adv_u(const lex_grid_t* g, const float* rho0_col, const float* u, const float* v, const float* w, float weps,
              const float* flow_div, const float* sfc_flux, float* out, void* stream);
/* advection_v (advection.py:77-114) -> (nx, ny+1, nz) */
int lex_adv_v(const lex_grid_t* g, const float* rho0_col, const float* u, const float* v, const float* w, float weps,
              const float* flow_div, const float* sfc_flux, float* out, void* stream);
/* advection_w (advection.py:117-149) -> (nx, ny, nz+1), zero at surface and lid */
int lex_adv_w(const lex_grid_t* g, const float* rho0_col, const float* u, const float* v, const float* w, float weps,
              const float* flow_div, float* out, void* stream);
/* advection_pi (advection.py:152-197) -> interior */
int lex_adv_pi(const lex_grid_t* g, const float* rho0_col, const float* pi0_col, const float* pip, const float* u,
               const float* v, const float* w, float* out, void* stream);

/* ---- B1: buoyancy --------------------------------------------------------------- */
/* calculate_buoyancy (pressure_equations.py:162-172) -> b interior, b8w (nx, ny, nz+1) */
int lex_buoyancy(const lex_grid_t* g, const float* theta0_col, const float* qv0_col, const float* theta,
                 const float* qv, float* b, float* b8w, void* stream);
/* density potential temperature theta(1+Rv/Rd qv)/(1+qv) (one_step_integration.py:135), n elements */
int lex_theta_rho(long n, const float* theta, const float* qv, float* out, void* stream);

/* ---- P1/P2: pressure gradient, Coriolis ----------------------------------------- */
/* pressure_gradient_force (pressure_gradient_coriolis.py:7-37) -> u, v, w interiors */
int lex_pgf(const lex_grid_t* g, const float* pip, const float* theta_rho, float* pu, float* pv, float* pw,
            void* stream);
/* calculate_coriolis_force (pressure_gradient_coriolis.py:80-95) -> fu (nx,ny+1,nz), fv (nx+1,ny,nz) */
int lex_coriolis(const lex_grid_t* g, const float* u, const float* v, float fcor, float* fu, float* fv, void* stream);
/* calculate_coriolis_force_with_lspgrad (pressure_gradient_coriolis.py:98-113):
 * fu = f (mean of u - u_geo), fv = f (mean of v - v_geo); the reference fixes
 * (u_geo, v_geo) = (1, 0) m/s */
int lex_coriolis_lspgrad(const lex_grid_t* g, const float* u, const float* v, float fcor, float u_geo, float v_geo,
                         float* fu, float* fv, void* stream);
/* the momentum right-hand sides of a stage (one_step_integration.py:126-133):
 * rhs_u = adv4u + fv + sgs_u (u points), rhs_v = adv4v - fu + sgs_v (v points),
 * rhs_w = adv4w + b8w + sgs_w (w points); fu/fv, b8w and the sgs terms may be NULL */
int lex_momentum_rhs(const lex_grid_t* g, const float* adv4u, const float* adv4v, const float* adv4w,
                     const float* fu, const float* fv, const float* b8w, const float* sgs_u, const float* sgs_v,
                     const float* sgs_w, float* rhs_u, float* rhs_v, float* rhs_w, void* stream);

/* ---- P3/P4: pressure equation ---------------------------------------------------- */
/* calculate_rtt (pressure_equations.py:175-195); buoyancy interior -> padded rtt */
int lex_rtt(const lex_grid_t* g, const float* rho0_theta0_col, const float* theta, const float* qv,
            const float* buoyancy, float* rtt, void* stream);
/* rhs_of_pressure_equation (pressure_equations.py:51-77); fu/fv may be NULL (zero) */
int lex_p_rhs(const lex_grid_t* g, const float* rho0_theta0_col, const float* adv4u, const float* adv4v,
              const float* adv4w, const float* fu, const float* fv, const float* buoyancy, float* rhs, void* stream);
/* laplace_of_pressure (pressure_equations.py:8-48), affine */
int lex_p_matvec(const lex_grid_t* g, const float* rtt, const float* x, float* out, void* stream);
/* jax.scipy.sparse.linalg.bicgstab as called by solve_pres_eqn
 * (one_step_integration.py:468-487): FP64 reductions, device-side loop control.
 * *k_out receives the iteration count or -10/-11 (breakdown); k_out may be NULL
 * (then the call does not synchronise the stream at the end). */
size_t lex_bicgstab_workspace_bytes(const lex_grid_t* g);
int lex_bicgstab(const lex_grid_t* g, const float* rtt, const float* rhs, const float* x0, float* x, double tol,
                 double atol, int maxiter, void* workspace, size_t workspace_bytes, int* k_out, void* stream);
/* lex_bicgstab without a workspace argument and without any device allocation: no
 * counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * loop state, partial sums and Krylov vectors live in a static device array of the
 * library (one per GPU, about 32 MiB, held until the process ends), so grids are limited
 * to nx*ny*nz <= LEX_BICGSTAB_NOALLOC_MAX_CELLS (larger ones return -3).  Results equal
 * lex_bicgstab's bit for bit (same kernels, same launch shape).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two streams, two host threads)
 * race and must be serialised by the caller. */
#define LEX_BICGSTAB_NOALLOC_MAX_CELLS 1048576
int lex_bicgstab_noalloc(const lex_grid_t* g, const float* rtt, const float* rhs, const float* x0, float* x,
                         double tol, double atol, int maxiter, int* k_out, void* stream);

/* ---- S1/M1/E1: acoustic substep and Euler updates --------------------------------- */
/* one pass of the loop in sound_rk_step0/1/2 (one_step_integration.py:528-553):
 * (u,v,w,pip) -> (u_out,...); rhs_* interior; pip_star, theta_rho padded */
int lex_acoustic_substep(const lex_grid_t* g, const float* rho0_col, const float* pi0_col, const float* u,
                         const float* v, const float* w, const float* pip, const float* pip_star, const float* rhs_u,
                         const float* rhs_v, const float* rhs_w, const float* theta_rho, float dt_sound, float* u_out,
                         float* v_out, float* w_out, float* pip_out, void* stream);
/* two consecutive passes of that loop (one_step_integration.py:528-553) in one launch,
 * the same arguments; results equal two lex_acoustic_substep calls.  nz <= 256. */
int lex_acoustic_substep_pair(const lex_grid_t* g, const float* rho0_col, const float* pi0_col, const float* u,
                              const float* v, const float* w, const float* pip, const float* pip_star,
                              const float* rhs_u, const float* rhs_v, const float* rhs_w, const float* theta_rho,
                              float dt_sound, float* u_out, float* v_out, float* w_out, float* pip_out, void* stream);
/* one pass of the loop in sound_implicit_rk_step0/1/2 (one_step_integration.py:590-633,
 * solver_opt 3): the same arguments plus the off-centring alpha (namelist alpha = 0.6);
 * w and pi' implicit in the vertical, one tridiagonal solve per column. */
int lex_acoustic_substep_implicit(const lex_grid_t* g, const float* rho0_col, const float* pi0_col, const float* u,
                                  const float* v, const float* w, const float* pip, const float* pip_star,
                                  const float* rhs_u, const float* rhs_v, const float* rhs_w, const float* theta_rho,
                                  float dt_sound, float alpha, float* u_out, float* v_out, float* w_out,
                                  float* pip_out, void* stream);
/* update_theta_euler / update_qv_euler (one_step_integration.py:450-465) plus the SGS
 * add-in (:158-162): out = pad(s0 + (adv+heating)*dt) + pad(sgs*dt).  heating, sgs,
 * sfc_flux and tend_out may be NULL. */
int lex_scalar_update(const lex_grid_t* g, const float* rho0_col, const float* s0, const float* s, const float* u,
                      const float* v, const float* w, const float* flow_div, const float* sfc_flux,
                      const float* heating, const float* sgs, float dt, float weps, float* out, float* tend_out,
                      void* stream);
/* update_momentum_eqn_euler (one_step_integration.py:876-891) plus SGS add-in
 * (:113-118); fu, fv, sgs_* may be NULL; du/dv/dw_out (interior tendencies incl. SGS,
 * as rk_sub_step0 returns them) may be NULL */
int lex_momentum_update(const lex_grid_t* g, const float* u0, const float* v0, const float* w0, const float* pip,
                        const float* theta_rho, const float* adv4u, const float* adv4v, const float* adv4w,
                        const float* fu, const float* fv, const float* b8w, const float* sgs_u, const float* sgs_v,
                        const float* sgs_w, float dt, float* u_out, float* v_out, float* w_out, float* du_out,
                        float* dv_out, float* dw_out, void* stream);

/* ---- G1: Smagorinsky -------------------------------------------------------------- */
/* compute_deformation (turbulence.py:298-388) -> 7 padded arrays */
int lex_smag_deformation(const lex_grid_t* g, const float* rho0_col, const float* u, const float* v, const float* w,
                         float* s11, float* s22, float* s33, float* s12, float* s13, float* s23, float* deform,
                         void* stream);
/* compute_nm (turbulence.py:391-427) -> padded n2 */
int lex_smag_nm(const lex_grid_t* g, const float* theta, const float* qv, float* n2, void* stream);
/* compute_k (turbulence.py:430-446); cs = 0.18 and prandtl = 1/3 in the reference */
int lex_smag_k(const lex_grid_t* g, const float* deform, const float* n2, float cs, float prandtl, float* km,
               float* kh, void* stream);
/* compute_smag (turbulence.py:9-111): momentum part -> sgs_u, sgs_v, sgs_w */
int lex_smag_momentum(const lex_grid_t* g, const float* rho0_col, const float* km, const float* s11,
                      const float* s22, const float* s33, const float* s12, const float* s13, const float* s23,
                      float* sgs_u, float* sgs_v, float* sgs_w, void* stream);
/* compute_smag (turbulence.py:80-109): one scalar's SGS tendency (interior) */
int lex_smag_scalar(const lex_grid_t* g, const float* rho0_col, const float* kh, const float* s, float* out,
                    void* stream);
/* compute_k2 (turbulence.py:449-471), turb_opt 2: horizontal Kmh = cs^2 dx dy Dh and vertical
 * Kmv = cs^2 Lv^2 sqrt(max(Dv - n2, 0)) (prandtl 1, so Khh = Kmh, Khv = Kmv) from the
 * deformation's s11..s23 and n2 (padded) -> padded kmh, kmv */
int lex_smag_k2(const lex_grid_t* g, const float* rho0_col, const float* s11, const float* s22, const float* s33,
                const float* s12, const float* s13, const float* s23, const float* n2, float cs, float* kmh,
                float* kmv, void* stream);
/* compute_smag2 (turbulence.py:114-295): momentum part with kmh on tau11, tau22, tau33,
 * tau12 and kmv on tau13, tau23.  sgs_w at the surface is written as zero (compute_smag2
 * leaves it at its stencil value; the step sets w to zero there after every substep). */
int lex_smag2_momentum(const lex_grid_t* g, const float* rho0_col, const float* kmh, const float* kmv,
                       const float* s11, const float* s22, const float* s33, const float* s12, const float* s13,
                       const float* s23, float* sgs_u, float* sgs_v, float* sgs_w, void* stream);
/* compute_smag2 (turbulence.py:207-234): one scalar's SGS tendency with khh on the x / y
 * faces and khv on the z faces */
int lex_smag2_scalar(const lex_grid_t* g, const float* rho0_col, const float* khh, const float* khv, const float* s,
                     float* out, void* stream);

/* ---- I1: hydrostatic base state (FP64 on the GPU) ---------------------------------- */
/* setup_ic_option1 (setup_lex.py:103-137): rh_col (nz+2 doubles, device) ->
 * out (5*(nz+2)+1 doubles, device): pi0, rho0_theta0, rho0, theta0, qv0, then the
 * iteration count; *iters (host, may be NULL) receives it after a stream sync. */
int lex_base_state(int nz, double dz, double theta0, const double* rh_col, double* out, int max_iter, double tol,
                   int* iters, void* stream);

/* ---- R1/R2: fused WS-RK3 step (solver_opt 2, compressible) ----------------------- */
typedef struct lex_step_opts {
  float dt;          /* full step (namelist dt) */
  int n_sound;       /* acoustic substeps per full step (namelist n_sound) */
  int turb;          /* 1 = Smagorinsky (turb_opt 1), 2 = horizontal / vertical Smagorinsky
                        (turb_opt 2, turbulence.py:114-295, 449-471), 0 = off */
  float cs;          /* Smagorinsky constant (turbulence.py:433) */
  float prandtl;     /* turbulence.py:432 */
  int n_passive;     /* passive scalars advected and diffused like qv (qc, qr) */
  int impl;          /* 0 = fused stage kernels (default), 1 = one kernel per reference routine */
  int solver;        /* acoustic substeps: 0 or 2 = forward-backward (solver_opt 2), 3 = vertically
                        implicit w / pi' (solver_opt 3, one_step_integration.py:559-873) */
  int damp;          /* 1 = upper-level Rayleigh damping after every stage (damp_opt,
                        boundary_conditions.py:119-141, one_step_integration.py:164-175) */
  int rad;           /* 1 = Newtonian heating of theta (rad_opt, radiation.py:7-16), computed from
                        the stage-0 theta and kept for the step (one_step_integration.py:31-35) */
  float z_damping;   /* height (m) above which the damping acts (namelist z_damping, 9600) */
  float rd_alpha;    /* inverse e-folding time (1/s) of the damping (namelist rd_alpha, 1/120) */
  int sfc;           /* 1 = bulk atmosphere-ocean surface fluxes (sfc_opt, boundary_conditions.py:7-75)
                        from the stage-0 state, kept for the step (one_step_integration.py:37-52):
                        stress on the bottom u / v faces, sen/Cp and evaporation through the
                        bottom theta / qv faces; needs surface_t in the base state */
  int cor;           /* Coriolis (cor_opt, one_step_integration.py:83-89): 0 off, 1 f-plane
                        (pressure_gradient_coriolis.py:80-95), 2 f-plane with a large-scale
                        pressure gradient balancing the geostrophic wind (:98-113) */
  float fcor;        /* Coriolis parameter (namelist fCor, 1e-4 / s) */
  float u_geo, v_geo;/* geostrophic wind for cor 2 (the reference's constants: 1, 0 m/s) */
} lex_step_opts_t;

typedef struct lex_base_cols {
  const float* rho0;
  const float* theta0;
  const float* qv0;
  const float* pi0;
  const float* rho0_theta0;
  const float* surface_t;  /* (nx, ny) sea-surface temperature (K), base_state[5]; read only with sfc */
} lex_base_cols_t;

/* ws_rk3_big_step (one_sprint_functions.py:10-30) for solver_opt 2 and 3: advances the
 * state in place.  state[] = theta, u, v, w, pip, qv, then n_passive scalars (all
 * padded).  The adjoint entry points below reject damp, rad and sfc. */
size_t lex_rk3_workspace_bytes(const lex_grid_t* g, int n_passive);
int lex_rk3_step(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                 void* workspace, size_t workspace_bytes, void* stream);
/* Byte offsets inside the step's workspace of the SGS tendencies the last step used
 * (frozen over its stages, one_step_integration.py:57-76, :245): offsets[0..2] = sgs_u,
 * sgs_v, sgs_w (interior, (nx+1, ny, nz), (nx, ny+1, nz), (nx, ny, nz+1)), offsets[3..]
 * = sgs_theta, sgs_qv, then one per passive scalar ((nx, ny, nz)); 3 + 2 + n_passive
 * entries.  They hold the last step's values until the next step on the workspace. */
int lex_rk3_workspace_sgs_offsets(const lex_grid_t* g, int n_passive, size_t* offsets);

/* ---- measurement (no reference counterpart: replaces the wall-clock table of
 * lex_relay.py:65-75 with per-kernel device timing) ------------------------------ */
/* Bracket every kernel the step driver launches with a HIP event pair (pool of
 * capacity_pairs, default 8192 when <= 0).  Process-global, single-threaded use. */
int lex_profile_enable(int on, int capacity_pairs);
/* Synchronise, sum elapsed ms and launch counts per kernel kind into ms[0..n),
 * counts[0..n) and reset the pool; returns the number of kernel kinds. */
int lex_profile_read(double* ms, int* counts, int n);
const char* lex_profile_kernel_name(int id);

/* ---- adjoint: vector-Jacobian products of the step -------------------------------
 * The reference differentiates implicitly with jax.value_and_grad (lex_dl_train.py:114);
 * these are the hand-written reverses of the forward entry points above.  Every *_vjp
 * ACCUMULATES into its adjoint outputs (the caller zeroes them).  Adjoints of padded
 * inputs are returned per padded position (ghosts included); lex_fold reduces them onto
 * the interior. */

/* Reverse of padding3_array / padding3u_array / padding3v_array
 * (one_step_integration.py:950-985): add every ghost's adjoint onto its interior source
 * and clear the ghost.  mode 0 scalar, 1 u, 2 v, 3 w (ghost z levels of w included). */
int lex_fold(const lex_grid_t* g, int mode, float* adj, void* stream);

/* advection_scalar (advection.py:9-34) with flow_divergence = get_divergence(u, v, w)
 * (:200-226) and a prescribed surface flux: lam (nx,ny,nz) -> ls, lu, lv, lw (padded).
 * The advection reverses use no atomics: contributions several threads share go through
 * slot planes (18 for the scalar, 24 per momentum component, each the size of the launch
 * box) that a gather adds in a fixed order, so their results are bit-reproducible.  The
 * planes are taken from the stream's pool (hipMallocAsync) and returned on the stream;
 * -4 when the pool has no room.  The library sets the device's default pool to keep
 * freed memory (release threshold UINT64_MAX) so repeated calls do not map it afresh. */
int lex_adv_scalar_vjp(const lex_grid_t* g, const float* rho0, const float* s, const float* u, const float* v,
                       const float* w, const float* lam, float weps, float* ls, float* lu, float* lv, float* lw,
                       void* stream);
/* advection_u / _v / _w (advection.py:37-149), weps 1e-16, flow divergence recomputed
 * from (u, v, w): lam on the u, v or w points -> lu, lv, lw (padded). */
int lex_adv_u_vjp(const lex_grid_t* g, const float* rho0, const float* u, const float* v, const float* w,
                  const float* lam, float* lu, float* lv, float* lw, void* stream);
int lex_adv_v_vjp(const lex_grid_t* g, const float* rho0, const float* u, const float* v, const float* w,
                  const float* lam, float* lu, float* lv, float* lw, void* stream);
int lex_adv_w_vjp(const lex_grid_t* g, const float* rho0, const float* u, const float* v, const float* w,
                  const float* lam, float* lu, float* lv, float* lw, void* stream);
/* b8w of calculate_buoyancy (pressure_equations.py:123-172): lam (nx,ny,nz+1) -> lth, lqv. */
int lex_buoyancy_vjp(const lex_grid_t* g, const float* theta0, const float* qv0, const float* th, const float* qv,
                     const float* lam_b8w, float* lth, float* lqv, void* stream);
/* theta_rho = th (1 + Rv/Rd qv) / (1 + qv) elementwise (one_step_integration.py:135). */
int lex_theta_rho_vjp(long n, const float* th, const float* qv, const float* ltr, float* lth, float* lqv,
                      void* stream);
/* One acoustic substep (one_step_integration.py:528-553): the adjoints of the substep's
 * outputs luo, lvo, lwo, lpo (padded; folded in place) -> adjoints of its inputs lu, lv,
 * lw, lp (u, v, w, pi'), of pi* (lps), theta_rho (ltr) and the rhs (lru, lrv, lrw).
 * lU, lV, lW: caller scratch of the interior u, v, w sizes. */
int lex_acoustic_substep_vjp(const lex_grid_t* g, const float* rho0, const float* pi0, const float* u,
                             const float* v, const float* w, const float* p, const float* ps, const float* ru,
                             const float* rv, const float* rw, const float* tr, float dts, float* luo, float* lvo,
                             float* lwo, float* lpo, float* lU, float* lV, float* lW, float* lu, float* lv,
                             float* lw, float* lp, float* lps, float* ltr, float* lru, float* lrv, float* lrw,
                             void* stream);
/* Smagorinsky chain compute_deformation -> compute_nm -> compute_k -> compute_smag
 * (turbulence.py:9-111, :298-446): adjoints of sgs_u, sgs_v, sgs_w and of each scalar's
 * SGS tendency -> lu, lv, lw, lth, lqv (padded), each scalar's adjoint, and dL/dCs
 * accumulated into *lcs (device FP64; may be NULL). */
size_t lex_smag_vjp_workspace_bytes(const lex_grid_t* g);
int lex_smag_vjp(const lex_grid_t* g, const float* rho0, const float* u, const float* v, const float* w,
                 const float* th, const float* qv, int n_scalars, const float* const* scalars, float cs,
                 float prandtl, const float* lsu, const float* lsv, const float* lsw, const float* const* lss,
                 float* lu, float* lv, float* lw, float* lth, float* lqv, float* const* lscalars, double* lcs,
                 void* workspace, size_t workspace_bytes, void* stream);
/* One WS-RK3 step (ws_rk3_big_step, one_sprint_functions.py:10-30; solver_opt 2):
 * re-runs the step with a tape, then reverses it.  lam_out: adjoint of the output state
 * (6 + n_passive padded fields; folded in place); lam_in: receives the adjoint of the
 * input state (folded: ghosts zero); lam_cs: dL/dCs accumulated (device FP64, may be
 * NULL).  state is not modified. */
size_t lex_rk3_vjp_workspace_bytes(const lex_grid_t* g, float dt, int n_sound, int n_passive);
int lex_rk3_step_vjp(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                     float* const* lam_out, float* const* lam_in, double* lam_cs, void* workspace,
                     size_t workspace_bytes, void* stream);
/* The same step split in two, so a checkpointed backward pass records each step of a
 * segment once and reverses it from the record: lex_rk3_step_tape advances state into
 * state_out (not aliased) recording the tape; lex_rk3_step_vjp_taped reverses that step
 * from its tape (the tape is only read).  The tape holds the stage states, the frozen
 * SGS fields, rhs and theta_rho per stage and every acoustic substep input. */
size_t lex_rk3_tape_bytes(const lex_grid_t* g, float dt, int n_sound, int n_passive);
size_t lex_rk3_adjoint_bytes(const lex_grid_t* g, float dt, int n_sound, int n_passive);
int lex_rk3_step_tape(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                      float* const* state_out, void* tape, size_t tape_bytes, void* stream);
int lex_rk3_step_vjp_taped(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                           float* const* state, const void* tape, size_t tape_bytes, float* const* lam_out,
                           float* const* lam_in, double* lam_cs, void* adjoint_ws, size_t adjoint_bytes,
                           void* stream);

/* ---- learned SGS correction net: 3x3x3 'SAME' convolutions on BF16 MFMA -----------
 * Replaces flax nn.Conv(kernel (3,3,3), strides 1, padding 'SAME', bias) and nn.gelu of
 * simpleCNN (dl_models.py:19-30), and their jax.grad.  Activations are channels-last
 * (x, y, z, c) BF16 with the channel count padded to 16 (c <= 16) or 32; weights FP32
 * W[kx*9 + ky*3 + kz][cin][cout] (flax's (3,3,3,in,out) kernel) packed to BF16 by
 * lex_conv3d_pack_weights (flip = 1 packs the transposed convolution of the data
 * gradient).  Accumulation is FP32.  Channels 1..64 (padded to 16, 32 or 64); the weight
 * gradient takes 1..32. */
size_t lex_conv3d_packed_weight_bytes(int cin, int cout, int flip);
int lex_conv3d_pack_weights(const float* w, int cin, int cout, int flip, void* wp, void* stream);
/* y = gelu?(conv(x) + bias): BF16 y (padded cout) or FP32 yf (cout channels); pre
 * (optional) receives the FP32 pre-activation (padded cout channels). */
int lex_conv3d_bf16_fwd(int nx, int ny, int nz, int cin, int cout, const void* x, const void* wp, const float* bias,
                        int gelu, int out_fp32, void* y, float* yf, float* pre, void* stream);
/* The same with periodic x and y and zero z padding: Zcnn's layers (dl_models.py:71-122),
 * flax padding 'CIRCULAR' on a column whose z ends are held at zero. */
int lex_conv3d_bf16_fwd_circular(int nx, int ny, int nz, int cin, int cout, const void* x, const void* wp,
                                 const float* bias, int gelu, int out_fp32, void* y, float* yf, float* pre,
                                 void* stream);
/* dx (FP32, cin channels) = transposed convolution of g_pre (BF16, padded cout
 * channels) with the flip-packed weights; zero_bias: padded-cin zeros. */
int lex_conv3d_bf16_bwd_data(int nx, int ny, int nz, int cin, int cout, const void* g_pre, const void* wp_flip,
                             const float* zero_bias, float* dx, void* stream);
/* The data gradient of a layer whose input came out of a GELU, and the GELU's backward
 * (jax.grad through the nn.gelu calls of simpleCNN, dl_models.py:24-28) in one pass:
 * g_pre_in (BF16, padded cin channels) = BF16(dx * gelu'(pre_in)), dx as
 * lex_conv3d_bf16_bwd_data computes it and pre_in the FP32 pre-activation of the layer
 * below (padded cin channels, as lex_conv3d_bf16_fwd writes it).  Bit for bit the same
 * as lex_conv3d_bf16_bwd_data followed by lex_gelu_bwd.  cout <= 32 and cin 17..32 (-1
 * otherwise: run the two calls). */
int lex_conv3d_bf16_bwd_data_gelu(int nx, int ny, int nz, int cin, int cout, const void* g_pre, const void* wp_flip,
                                  const float* zero_bias, const float* pre_in, void* g_pre_in, void* stream);
/* dw[27][cin_p][cout_p] += sum_p x[p + o] g_pre[p]; db[cout] += sum_p g_pre[p] (db may
 * be NULL).  x: BF16, padded cin; g_pre: BF16, padded cout.  Bit-reproducible: per-
 * workgroup partial sums (stream-ordered scratch from the stream's pool, returned on the
 * stream; -4 when the pool has no room) are added in workgroup order. */
int lex_conv3d_bf16_bwd_weight(int nx, int ny, int nz, int cin, int cout, const void* x, const void* g_pre, float* dw,
                               float* db, void* stream);
/* AutoEncoder (dl_models.py:33-68) layers in FP32, channels last: one correlation over
 * the input dilated by lhs_dilation and zero-padded (pad_lo per axis; the high side is
 * implied by the output extent), window stride `stride`, kernel k^3 (k = 3, or 1 for the
 * dense layer) laid out [tap][cin][cout] like flax's kernels; bias may be NULL; relu 1
 * applies max(x, 0).  flax Conv(stride s, 'SAME') = (stride s, dilation 1, pad_lo =
 * total // 2); flax ConvTranspose(stride s) = (stride 1, dilation s, lax.conv_transpose's
 * padding); nn.Dense = k 1 on the flattened input. */
int lex_conv3d_f32(const float* x, int nix, int niy, int niz, int cin, const float* w, const float* bias, int k,
                   int stride, int lhs_dilation, int pad_lo_x, int pad_lo_y, int pad_lo_z, int relu, float* y, int nox,
                   int noy, int noz, int cout, void* stream);
/* g_pre = g_post * gelu'(pre) (tanh form) or g_post when gelu == 0, as BF16 with cp
 * channels (channels >= c_real zero). */
int lex_gelu_bwd(long ncell, int cp, int c_real, int cpre, const float* gpost, int gstride, const float* pre, int gelu,
                 void* gpre, void* stream);
/* FP32 channels-last (c_real channels at stride cstride) -> BF16 with cp channels. */
int lex_to_bf16_channels_last(long ncell, int c_real, int cstride, int cp, const float* in, void* out, void* stream);

/* Net inputs: 11 cell-centred channels (u, v, w at centres, theta - theta0, qv - qv0,
 * and S11, S22, S33, S12, S13, S23 staggered as compute_deformation,
 * turbulence.py:298-388, without the rho factor), out = (x - shift[c]) * scale[c] (shift,
 * scale: host arrays of 11),
 * channels-last FP32 (nx, ny, nz, 11).  The reference's packing of its own 5 inputs is
 * phys_state_tuple2array (dl_models.py:125-155). */
int lex_sgs_features(const lex_grid_t* g, const float* theta0, const float* qv0, const float* u, const float* v,
                     const float* w, const float* th, const float* qv, const float* shift, const float* scale,
                     float* out, void* stream);
/* Net outputs y (nx, ny, nz, 5), times out_scale[c] (host array of 5), -> SGS tendencies
 * on compute_smag's points (turbulence.py:9-111): sgs_u (nx+1,ny,nz), sgs_v
 * (nx,ny+1,nz), sgs_w (nx,ny,nz+1; zero at the lids), sgs_theta, sgs_qv (nx,ny,nz). */
int lex_sgs_apply(const lex_grid_t* g, const float* y, const float* out_scale, float* su, float* sv, float* sw,
                  float* sth, float* sqv, void* stream);
/* Block-mean coarsening by (fx, fy, fz) (initial_comparison/plot_theta_pert_coarsened.m:7-8)
 * of a padded fine field into an interior coarse field: mode 0 cell mean of a scalar,
 * 1/2/3 face-area means of u/v/w on the coarse faces. */
int lex_coarsen(const lex_grid_t* g, int mode, int fx, int fy, int fz, const float* a, float* out, void* stream);
/* The step with the SGS tendencies recomputed at the start of every WS-RK3 stage from
 * that stage's input state (the learned SGS net run inside each stage); no counterpart
 * in the reference, whose SGS is computed in stage 0 and frozen
 * (one_step_integration.py:57-76).  fn(stage, stage_state, su, sv, sw, ss, user) is called
 * on the host thread for stage = 0, 1, 2 with the stage's 6 + n_passive padded state
 * fields and the library's SGS slots (u, v, w points; scalars interior); it fills them by
 * work it enqueues on `stream` and returns 0 (non-zero aborts the step with -5). */
typedef int (*lex_stage_sgs_fn)(int stage, float* const* stage_state, float* sgs_u, float* sgs_v, float* sgs_w,
                                float* const* sgs_scalars, void* user);
int lex_rk3_step_stage_sgs(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                           float* const* state, lex_stage_sgs_fn fn, void* user, void* workspace,
                           size_t workspace_bytes, void* stream);
/* lex_rk3_step with caller-supplied SGS tendencies (frozen over the step, as the
 * stage-0 Smagorinsky tendencies are, one_step_integration.py:57-76): su, sv, sw on the
 * u, v, w points, ss[] for theta, qv and the passive scalars.  o->turb is ignored. */
int lex_rk3_step_ext_sgs(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b, float* const* state,
                         const float* su, const float* sv, const float* sw, const float* const* ss, void* workspace,
                         size_t workspace_bytes, void* stream);

/* ---- training the net through the solver ---------------------------------------------
 * Replace jax.value_and_grad of the loss through ws_rk3_sprint_dl_train in train_step
 * (lex_dl_train.py:111-131; the sprint is one_sprint_functions.py:118-165) for the
 * build's in-stage SGS net.  Reverses of lex_sgs_apply (ly written) and of
 * lex_sgs_features (padded adjoints of u, v, w, theta, qv accumulated, ghost lines
 * included: fold with lex_fold afterwards; scale as passed to lex_sgs_features; gathered
 * per point, no atomics). */
int lex_sgs_apply_vjp(const lex_grid_t* g, const float* lsu, const float* lsv, const float* lsw, const float* lsth,
                      const float* lsqv, const float* out_scale, float* ly, void* stream);
int lex_sgs_features_vjp(const lex_grid_t* g, const float* lfeat, const float* scale, float* lu, float* lv, float* lw,
                         float* lth, float* lqv, void* stream);
/* lex_rk3_step_tape / lex_rk3_step_vjp_taped with caller-supplied SGS tendencies (as
 * lex_rk3_step_ext_sgs): the reverse writes the adjoints of su, sv, sw and ss[] (same
 * shapes) instead of taking them through the Smagorinsky reverse. */
int lex_rk3_step_tape_ext_sgs(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                              float* const* state, const float* su, const float* sv, const float* sw,
                              const float* const* ss, float* const* state_out, void* tape, size_t tape_bytes,
                              void* stream);
int lex_rk3_step_vjp_taped_ext_sgs(const lex_grid_t* g, const lex_step_opts_t* o, const lex_base_cols_t* b,
                                   float* const* state, const void* tape, size_t tape_bytes, float* const* lam_out,
                                   float* const* lam_in, float* lsu, float* lsv, float* lsw, float* const* lss,
                                   void* adjoint_ws, size_t adjoint_bytes, void* stream);

/* ---- the reference's post-step deep-learning correction ------------------------------
 * The pieces of ws_rk3_sprint_dl (one_sprint_functions.py:118-165) around its network:
 * phys_state_tuple2array (dl_models.py:125-155): padded theta, u, v, w, qv -> out
 * (nx, ny, nz, 5) channels-last FP32, each field normalised per level as (x - lo) / den
 * (lo: min over x, y of the scaling minimum; den: max over x, y of the scaling maximum
 * - lo + 0.01, or + 1e-5 for qv), u, v, w averaged to centres.  lo and den concatenate
 * the per-level values of theta, u, v (nz each), w (nz+1) and qv (nz): 5 nz + 1 floats. */
int lex_dl_pack(const lex_grid_t* g, const float* theta, const float* u, const float* v, const float* w,
                const float* qv, const float* lo, const float* den, float* out, void* stream);
/* correct_array2tuple with stagger_uvw (dl_models.py:158-182): the net's (nx, ny, nz, 5)
 * centred output -> corrections on theta (nx,ny,nz), u (nx+1,ny,nz), v (nx,ny+1,nz),
 * w (nx,ny,nz+1), qv points, with the reference's end faces. */
int lex_dl_stagger(const lex_grid_t* g, const float* c, float* theta_c, float* u_c, float* v_c, float* w_c,
                   float* qv_c, void* stream);
/* The correction step of ws_rk3_sprint_dl (one_sprint_functions.py:137-160): padded
 * theta, u, v, w, qv += staggered correction x range (r_*: the scaling maximum - minimum
 * on each field's points, interior shapes), in place, ghosts refilled; pi' untouched. */
int lex_dl_apply(const lex_grid_t* g, const float* c, const float* r_theta, const float* r_u, const float* r_v,
                 const float* r_w, const float* r_qv, float* theta, float* u, float* v, float* w, float* qv,
                 void* stream);

/* ---- the loss of the post-step correction's training (lex_dl_train.py:49-99) ---------
 * laplace_of_tensor (dataset_generation.py:271-320) of nfields stacked interior scalar
 * fields (nx, ny, nz): periodic x, y; end levels copied in z; no vertical term in the first
 * and last x planes (the reference's s_z[:, :, (0, -1)] = 0 on 5-D stacks indexes x). */
int lex_laplace(const lex_grid_t* g, long nfields, const float* s, float* out, void* stream);
/* Deterministic FP64 sum over n elements (*out, device) of op 0: w * 0.5 (a - b)^2
 * (optax.l2_loss times a weight) or op 1: w * max(-a, 0) (the negative-qv penalty), with
 * w repeating every `period` elements (b unused for op 1).  Workspace: device memory of
 * lex_weighted_sum_workspace_bytes(). */
/* The gradient of those terms with respect to a, accumulated into out (n floats):
 * op 0: w (a - b); op 1: -w where a < 0. */
int lex_weighted_sum_grad(long n, int op, const float* a, const float* b, const float* w, long period, float* out,
                          void* stream);
/* Reverses for the training's gradient through the post-step correction
 * (one_sprint_functions.py:54-83, differentiated by jax.value_and_grad in the reference,
 * lex_dl_train.py:114): lex_dl_apply_vjp writes dc (nx, ny, nz, 5), the adjoint of the
 * network's output, from the adjoint of the corrected state (padded, ghosts folded) and
 * the ranges of lex_dl_apply; lex_dl_pack_vjp ADDS the adjoint of the state's interior
 * values for the adjoint dp (nx, ny, nz, 5) of lex_dl_pack's output (lo unused: the map
 * is affine). */
int lex_dl_apply_vjp(const lex_grid_t* g, const float* l_theta, const float* l_u, const float* l_v, const float* l_w,
                     const float* l_qv, const float* r_theta, const float* r_u, const float* r_v, const float* r_w,
                     const float* r_qv, float* dc, void* stream);
int lex_dl_pack_vjp(const lex_grid_t* g, const float* dp, const float* lo, const float* den, float* l_theta,
                    float* l_u, float* l_v, float* l_w, float* l_qv, void* stream);
size_t lex_weighted_sum_workspace_bytes(void);
int lex_weighted_sum(long n, int op, const float* a, const float* b, const float* w, long period, double* out,
                     void* workspace, size_t workspace_bytes, void* stream);

/* ---- surface fluxes (sfc_opt) ----------------------------------------------------------
 * atm_ocn_flux (boundary_conditions.py:7-75) on n columns of device FP32 inputs (the
 * reference's arguments: z_bottom, u_bottom, v_bottom, theta_bottom, q_bottom,
 * rho_bottom, surface_t) -> out[0..6] = tau_x, tau_y, sen, evap, t_ref, q_ref, u10n
 * (n floats each, device), with the reference's signs (upward fluxes positive). */
int lex_atm_ocn_flux(long n, const float* z_bottom, const float* u_bottom, const float* v_bottom,
                     const float* theta_bottom, const float* q_bottom, const float* rho_bottom,
                     const float* surface_t, float* const* out, void* stream);
/* The same on the lowest level of a padded state as rk_sub_step0 slices it
 * (one_step_integration.py:37-47): u, v averaged to the cell centre, theta, qv and
 * rho0 (column, nz+2) of level 1, z_bottom = dz/2; out[0..6] (nx, ny) as above. */
int lex_surface_fluxes(const lex_grid_t* g, const float* rho0, const float* theta, const float* u, const float* v,
                       const float* qv, const float* surface_t, float* const* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* LEX_CAPI_H */
