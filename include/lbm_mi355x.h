/*
 * lbm_mi355x.h -- C ABI of the MI355X sparse D3Q19 MRT library (liblbm_mi355x.so).
 *
 * Drop-in boundary for the hot path of the jianhui-lbm porous-media solvers.  The reference
 * has no library interface: its path is a set of free C++ functions over global state inside
 * two programs,
 *   SP = Single_Phase_3D/SINGLE_PHASE_MPI_Software_Spars.cpp
 *   CG = MP_3D/Multi_Component_CG_MPI_SPARSE_Software.cpp
 * Each entry point below names the reference routine(s) it replaces (file:line).
 *
 * Conventions
 *   - Every function returns LBM_OK (0) or a negative LBM_ERR_* code; nothing exits.
 *     lbm_last_error() gives a message for the most recent failure on a handle.
 *   - A handle owns all of its device memory.  Host pointers are borrowed for the duration
 *     of the call only.  One host thread per handle at a time.
 *   - All work is ordered on the handle's own HIP stream.  Calls that return data to the host
 *     (get_state, diagnostics) synchronise that stream.
 *   - Geometry: int32 solid flags indexed [i][j][k] (z fastest), 0 = pore, non-zero = solid,
 *     exactly the reference's binary geometry file (SP:1058-1065).  The box is periodic in
 *     x, y and z (the reference's x-slab ring closes on itself, SP:1954-1999).
 *   - Fluid nodes keep the reference's numbering: i-major (x slowest, z fastest) over pore
 *     voxels (SP:1172-1193), 0-based here (the reference is 1-based with a padding record).
 *   - Host state arrays are FP64 in the reference's layouts: f[count][19] in the solver's own
 *     velocity ordering (SP:157-159 for single phase, CG:203-205 for colour gradient),
 *     u[count][3], scalars [count].
 */
#ifndef LBM_MI355X_H
#define LBM_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes */
#define LBM_OK 0
#define LBM_ERR_ARG (-22)       /* invalid argument */
#define LBM_ERR_NOMEM (-12)     /* device or host allocation failed */
#define LBM_ERR_HIP (-5)        /* HIP runtime error (message has the detail) */
#define LBM_ERR_STATE (-71)     /* call not valid in the handle's current state */
#define LBM_ERR_BUSY (-16)      /* serving: no room for the request right now */

/* ------------------------------------------------------------------ precision */
#define LBM_FP64 0   /* populations stored and collided in FP64 (parity build) */
#define LBM_FP32 1   /* FP32 populations stored as f_i - w_i*rho0, FP32 collide */
#define LBM_FP16 2   /* FP16 storage of (f_i - w_i*rho0)*scale, FP32 collide */

typedef struct lbm_handle lbm_handle;

/* ------------------------------------------------------------------ single phase */
typedef struct lbm_sp_params {
    int32_t nx, ny, nz;   /* box size in voxels (the counts of INPUT.dat line 3) */
    double nu;            /* kinematic viscosity in_vis (SP:1301): tau = 3 nu + 1/2 */
    double g[3];          /* body force gx, gy, gz (Guo forcing, SP:1519-1593) */
    double rho0;          /* rest density of the shifted storage, 1.0 in the reference */
    int32_t precision;    /* LBM_FP64 / LBM_FP32 / LBM_FP16 */
    double fp16_scale;    /* LBM_FP16 only: storage scale (power of two); 0 = default 4096 */
    int32_t device;       /* HIP device ordinal */
} lbm_sp_params;

typedef struct lbm_sp_diag {
    double Q[3];        /* sum of u over fluid nodes (in the box if given) -- Comput_Perm, SP:4617-4645 */
    double err;         /* sqrt(sum|u-u0|^2)/(sqrt(sum|u|^2)+1e-30), last node skipped -- Error, SP:3651-3664 */
    double umax;        /* max |u| over the same nodes (SP:3660-3665) */
    double sum_speed;   /* sum |u| over the same nodes (SP:3659), u_ave = sum_speed/(V*porosity) */
    double nan_count;   /* number of nodes whose u is not finite (the reference's NaN stop, SP:838) */
    double count_in_box;/* fluid nodes inside the box (all nodes if no box) */
} lbm_sp_diag;

/* Geometry, sparse numbering, neighbour lists and device allocation.
 * Replaces Parallelize_Geometry + init_Sparse_read_rock_parallel + Suppliment + the
 * allocations of main (SP:972-1145, SP:1147-1277, SP:486-534). */
int lbm_sp_create(const lbm_sp_params *p, const int32_t *solid_ijk, lbm_handle **out);
/* number of fluid nodes (Count) */
int lbm_sp_count(const lbm_handle *h, int64_t *count);

/* Pressure and velocity faces (the INPUT.dat lines 6-11 and Sub_BC, SP:262-267, 280).  Arrays
 * are indexed x+, x-, y+, y-, z+, z-.  After each streaming step the populations of the fluid
 * nodes on an active face are replaced as boundary_pressure (SP:3076-3620) and then
 * boundary_velocity (SP:2751-3075) do, faces in the reference's order y+, y-, z+, z-, x+, x-:
 *   pressure Sub_BC 0: f = feq(p, 0);  1: f = feq(p, u_nb);  2: f = M^-1 (M f_nb with m0 = p);
 *            3: f = feq(p, u_nb) + f_nb - feq(rho_nb, u_nb);  (nb = the inward neighbour, its
 *            pre-collision state of the same step; feq(p, 0) when nb is solid)
 *   velocity Sub_BC 0, 1: f = feq(1, v n)  (n the face's axis)
 * with the reference's two quirks (the x- face's u_ls[2] = u_nb[1] in Sub_BC 1 and 3, and no
 * replacement on the x faces under Sub_BC 2 where nb is solid).  Velocity faces with Sub_BC 2
 * or 3 are refused (LBM_ERR_ARG).  The streaming itself stays periodic, as in the reference.
 * Call before stepping (after create, init or load_state); NULL removes the faces. */
typedef struct lbm_sp_bc {
    int32_t sub_bc;
    int32_t pre[6];      /* pressure face on (1) / off (0): pre_xp, pre_xn, pre_yp, ... */
    double p[6];         /* boundary density p_xp, p_xn, ... */
    int32_t vel[6];      /* velocity face on / off: vel_xp, vel_xn, ... */
    double v[6];         /* face-normal velocity v_xp, v_xn, ... */
} lbm_sp_bc;
int lbm_sp_set_boundaries(lbm_handle *h, const lbm_sp_bc *bc);

/* Power-law fluid (the INPUT's nnf_module = 1 with nnf_m, nnf_n; collision_nnf, SP:2017-2700).
 * From the 12th step after init / load_state (the reference's loop step n > 10, SP:2397) each
 * node relaxes with its own viscosity nu = m |S|^(n-1), |S| the norm of the strain rate that
 * the non-equilibrium populations give with the node's viscosity of the previous step
 * (SP:2399-2491); init and load_state set every node's viscosity to the handle's nu.  m <= 0
 * switches it off.  Call before stepping (after create, init or load_state). */
int lbm_sp_set_power_law(lbm_handle *h, double m, double n);
/* per-node viscosity niu_vn[count] (the handle's nu everywhere without the power law) */
int lbm_sp_get_viscosity(lbm_handle *h, double *nu);
/* Post-streaming u and rho on the dense box, [i][j][k] (u: [i][j][k][3]), solid voxels u = 0,
 * rho = 1: the arrays the reference gathers on rank 0 for its VTK and profile writers
 * (output_velocity_b SP:4265-4390, output_density_b SP:4405-4505), scattered on the device into
 * 16 MiB chunks of the box and copied back chunk by chunk.  Either pointer may be NULL. */
int lbm_sp_get_dense(lbm_handle *h, double *u_ijk, double *rho_ijk);
/* f = feq(rho, u0), u = u0, rho everywhere -- init (SP:1282-1383) */
int lbm_sp_init(lbm_handle *h, double rho, const double u0[3]);
/* load a post-streaming state, e.g. a reference checkpoint -- Backup_init (SP:4715-4828).
 * u and rho are used by the first collision as given, like the reference does. */
int lbm_sp_load_state(lbm_handle *h, const double *f, const double *u, const double *rho);
/* nsteps x (collision + comput_macro_variables), the loop body of SP:662-684
 * (collision SP:1406-2013, comput_macro_variables SP:2702-2746), fused into one
 * pull-stream + collide kernel launch per step */
int lbm_sp_step(lbm_handle *h, int32_t nsteps);
/* same, and report the device time of the nsteps launches (HIP events on the handle stream) */
int lbm_sp_step_timed(lbm_handle *h, int32_t nsteps, double *elapsed_ms);
/* post-streaming f, u and rho in the reference's layout -- the content of Backup (SP:4507-4562)
 * without the padding record.  Any pointer may be NULL. */
int lbm_sp_get_state(lbm_handle *h, double *f, double *u, double *rho);
/* lbm_sp_get_state with no device allocation: no counterpart in the reference (whose Backup writes
 * from host arrays).  For callers that cannot allocate device memory between steps -- serving
 * loops that checkpoint a live state, harnesses that forbid hipMalloc after set-up.  The state
 * crosses through a fixed 16 MiB device array of the library (one per GPU, held until the process
 * ends) in chunks of 91,136 nodes, so any node count is accepted.  Calls on one GPU share that
 * array: calls that overlap in time on one GPU race (one call at a time per GPU). */
int lbm_sp_get_state_noalloc(lbm_handle *h, double *f, double *u, double *rho);
/* FP64 device reductions for Error (SP:3625-3704) and Comput_Perm (SP:4565-4712).
 * box = {xn, xp, yn, yp, zn, zp} inclusive voxel bounds (par_per_* / per_*) or NULL. */
int lbm_sp_diagnostics(lbm_handle *h, const int32_t *box, lbm_sp_diag *out);
/* Error (SP:3625-3704) as a run of the reference on nranks MPI ranks computes it: the box split
 * into x-slabs over the ranks as Parallelize_Geometry splits it (SP:1085-1108), each rank's sums
 * over its own nodes but its last (SP:3651), err_r = sqrt(sum|u-u0|^2)/(sqrt(sum|u|^2)+1e-30);
 * out[0] = the largest err_r, out[1] = the largest per-rank u_max, out[2] = the sum of |u| over
 * the ranks (SP:3677-3685).  nranks = 1 is lbm_sp_diagnostics' err, umax and sum_speed. */
int lbm_sp_error_ranks(lbm_handle *h, int32_t nranks, double out[3]);
/* Comput_Grop_Perm (SP:4886-5090), one cube grid in one pass over its nodes: the m^3 cubes of
 * edge `size` voxels, cube (ix, iy, iz) spanning origin + (ix, iy, iz) * size to that + size - 1
 * inclusive, at index c = (ix * m + iy) * m + iz (the order of the reference's loops and of the
 * lines of Perm_<size>.output).  out[4c .. 4c+2] = the sum of u over the cube's fluid nodes,
 * out[4c+3] = their number; voxels outside the box hold none.  out holds 4 m^3 doubles;
 * 1 <= size, 1 <= m, m^3 <= 2^24.  A grid's node lists are built on the first call and kept
 * (the four most recent grids). */
int lbm_sp_cube_sums(lbm_handle *h, const int32_t origin[3], int32_t size, int32_t m, double *out);
/* number of time steps taken since create/init/load */
int lbm_sp_steps_done(const lbm_handle *h, int64_t *steps);

/* ------------------------------------------------------------------ colour gradient */
typedef struct lbm_cg_params {
    int32_t nx, ny, nz;       /* box size in voxels (INPUT_CG_SPARSE.dat line 4) */
    double nu_l, nu_g;        /* viscosities of component A (psi = 1) and B (psi = -1), lines 19-20 */
    double cos_theta;         /* ContactAngle_parameter: psi at solid sites (CG:2512) */
    double cap_a;             /* CapA, surface-tension parameter (CG:3137-3143) */
    double g[3];              /* body force gxs, gys, gzs */
    int32_t bodyforce_apply;  /* 0 both phases; 1 / -1 force only psi >= rel_perm_psi / psi <= -rel_perm_psi2 (CG:3040) */
    double rel_perm_psi;      /* phase-1 flux threshold, also the force gate (CG:6366-6404) */
    double rel_perm_psi2;     /* phase-2 flux threshold */
    double rho0;              /* rest density of the shifted storage, 1.0 */
    int32_t precision;        /* LBM_FP64 / LBM_FP32 (FP16 is not offered for the colour-gradient path) */
    int32_t device;
    int32_t perm_dir;         /* PerDir 1..3: flow axis of the saturation displacement sums */
} lbm_cg_params;

typedef struct lbm_cg_diag {
    double Ql[3];        /* sum u over psi >= rel_perm_psi (in the box) -- Comput_Perm, CG:6366-6404 */
    double Qg[3];        /* sum u over psi <= -rel_perm_psi2 */
    double err, umax, sum_speed, nan_count;   /* as lbm_sp_diag (Error, CG:4979-5040) */
    double wet_count;    /* nodes with psi >= 0 in the box -- Comput_Saturation, CG:6640-6678 */
    double count_in_box; /* fluid nodes in the box */
    double sd_l, sd_g;   /* sum of psi * (coordinate along perm_dir) over psi >= 0 / psi < 0 (CG:6655-6667) */
} lbm_cg_diag;

/* geometry, numbering, neighbour lists, allocation (CG:1887-2308, CG:598-660) */
int lbm_cg_create(const lbm_cg_params *p, const int32_t *solid_ijk, lbm_handle **out);
int lbm_cg_count(const lbm_handle *h, int64_t *count);
/* init from a phase field given per voxel, double [i][j][k], clamped to [-1, 1] as the
 * reference clamps phase.dat (CG:2150-2155); rho = 1, rho_r = (psi+1)/2, u = u0 -- the
 * ini_Sat < 0 path of init (CG:2514-2547) */
int lbm_cg_init(lbm_handle *h, const double *psi_ijk, const double u0[3]);
/* load a post-streaming state in the CG velocity ordering -- Backup_init (CG:6765-7018) */
int lbm_cg_load_state(lbm_handle *h, const double *f, const double *u, const double *rho, const double *psi,
                      const double *rho_r, const double *rho_b);
/* nsteps x (collision + comput_macro_variables) (CG:938, CG:966; collision CG:2687-3660,
 * comput_macro_variables CG:3698-4099): per step one colour-gradient collide launch and
 * one colour-streaming launch */
int lbm_cg_step(lbm_handle *h, int32_t nsteps);
int lbm_cg_step_timed(lbm_handle *h, int32_t nsteps, double *elapsed_ms);
/* change the body force between steps, e.g. the stabiliser switch at stab_time (CG:934-935) */
int lbm_cg_set_body_force(lbm_handle *h, const double g[3]);
/* state in the reference's layouts (CG ordering for f); any pointer may be NULL -- Backup (CG:7020-7090) */
int lbm_cg_get_state(lbm_handle *h, double *f, double *u, double *rho, double *psi, double *rho_r, double *rho_b);
/* FP64 device reductions: phase fluxes, Error, saturation count (CG:4979, 6298, 6604) */
int lbm_cg_diagnostics(lbm_handle *h, const int32_t *box, lbm_cg_diag *out);
int lbm_cg_steps_done(const lbm_handle *h, int64_t *steps);

/* ------------------------------------------------------------------ passive solute
 * Advection-diffusion of a passive solute on the fluid nodes: D3Q7 BGK with rest weight w0,
 * w_q = (1 - w0)/6 for the six axis directions, c_s^2 = (1 - w0)/3, g_eq_q = w_q C (1 + e_q.u/c_s^2),
 * tau_s = D/c_s^2 + 1/2, halfway bounce-back (zero flux) at solid sites, pull streaming.  With
 * w0 = 0 this is the reference's six-velocity program (Single_Phase_3D/
 * Solute_Transport_MPI_Software_Spars_6V.cpp, ST6 below: g_eq = C/6 (1 + 3 e.u), ST6:1306-1310,
 * 1466-1535, 1625-1640); w0 = 1/4 is the usual D3Q7.  The velocity is either a frozen field
 * given per node (the reference's Velocity_for_solute files, ST6:3640-3665) or, coupled, the
 * velocity a single-phase handle computes in the same step. */
typedef struct lbm_st_params {
    int32_t nx, ny, nz;       /* box (ignored when coupled: the flow handle's box is used) */
    double D;                 /* diffusion coefficient niu_s */
    double w0;                /* rest weight, 0 <= w0 < 1 */
    int32_t precision;        /* LBM_FP64 or LBM_FP32 (coupled: must match the flow handle's compute type) */
    int32_t device;
} lbm_st_params;

typedef struct lbm_st_diag {
    double mass;              /* sum of C over the fluid nodes (FP64 mass balance) */
    double cmin, cmax;
    double moment1[3];        /* sum C x_a over nodes (voxel coordinates, no periodic unwrapping) */
    double moment2[3];        /* sum C x_a^2 */
    double nan_count;
} lbm_st_diag;

/* standalone: own numbering and neighbour lists, frozen velocity (zero until set) */
int lbm_st_create(const lbm_st_params *p, const int32_t *solid_ijk, lbm_handle **out);
/* coupled to a single-phase handle: same nodes and neighbour lists on the device; each
 * lbm_st_step takes one flow step and one solute step with that step's velocity, on the flow
 * handle's stream.  The flow handle must outlive this one. */
int lbm_st_create_coupled(const lbm_st_params *p, lbm_handle *flow, lbm_handle **out);
int lbm_st_count(const lbm_handle *h, int64_t *count);
/* frozen velocity u[count][3] in node order -- Backup_input_v (ST6:3640-3665) */
int lbm_st_set_velocity(lbm_handle *h, const double *u);
/* concentration per voxel, double [i][j][k]; g = g_eq(C, u) -- init (ST6:1249-1320) */
int lbm_st_init(lbm_handle *h, const double *conc_ijk);
/* post-streaming state g[count][7] (index 0 = rest) -- Backup_init (ST6:3716-3770) */
int lbm_st_load_state(lbm_handle *h, const double *g);
/* nsteps x (collision + streaming + C), the loop of ST6:525-546 (coupled: a flow step each too) */
int lbm_st_step(lbm_handle *h, int32_t nsteps);
int lbm_st_step_timed(lbm_handle *h, int32_t nsteps, double *elapsed_ms);
/* post-streaming g[count][7] and C[count] (either may be NULL) -- Backup (ST6:3681-3714) */
int lbm_st_get_state(lbm_handle *h, double *g, double *conc);
int lbm_st_diagnostics(lbm_handle *h, lbm_st_diag *out);

/* ------------------------------------------------------------------ permeability serving
 * A persistent request loop for many independent single-phase permeability solves -- the
 * GPU counterpart of the reference's practice of cutting a rock into sub-samples and running
 * one solver per sub-sample INPUT file (CPP_SCRIPT/HPC_SUBDOMAIN_INPUT*.cpp).  Requests of
 * uneven sizes are admitted into one device arena by HBM footprint, stepped together (one
 * launch per time step over the tiles of every active sample, per-sample parameters on the
 * device), checked every check_interval steps with per-sample FP64 reductions, and retired
 * when |k_n - k_(n-interval)| <= tol |k_n|.  A sample's result does not depend on which other
 * samples share its launches.  Calls are thread-safe; the loop runs on its own host thread. */
typedef struct lbm_srv_params {
    int64_t capacity_nodes;   /* arena size in fluid nodes; 0 = mem_fraction of free device memory (at most
                                 64 GB); capped at 2^32 / (19 * value size) nodes, 56.5 M FP32, 28.2 M FP64 */
    double mem_fraction;      /* used when capacity_nodes == 0 (default 0.5) */
    int32_t check_interval;   /* steps between permeability checks (freRe), default 100 */
    double tol;               /* relative permeability change that retires a sample, default 1e-6 */
    int32_t min_steps;        /* never retire before this many steps (default 2*check_interval) */
    int32_t max_steps;        /* retire unconverged after this many steps (default 200000) */
    int32_t precision;        /* LBM_FP32 (default) or LBM_FP64; LBM_FP16 is refused (LBM_ERR_ARG) */
    int32_t block;            /* nodes per workgroup tile: 128, 256 or 512; 0 = autotune at create */
    int32_t device;
    /* Anderson acceleration of each sample's steady state (no counterpart in the reference,
     * which only time-steps): every accel_period steps the K-step map x -> g(x) of the sample
     * is extrapolated from its last accel_depth residuals, x <- g - dG gamma with gamma the
     * least-squares fit of the newest residual by the residual differences.  The sample then
     * retires when k changes by <= tol relative between two accelerations. */
    int32_t accel_depth;      /* history depth m, 0 = plain time stepping, at most 8 */
    int32_t accel_period;     /* steps between accelerations, a multiple of check_interval (0 = 2*check_interval) */
    int32_t accel_warmup;     /* plain steps before the first acceleration, multiple of check_interval (0 = 5*check_interval) */
    /* Precision escalation of an FP32 server: a sample whose k (lattice units) falls below
     * escalate_k after the warm-up, or that reaches max_steps unsettled, moves with its state to
     * an internal FP64 server and continues there (its flux is then too small for FP32
     * deviations to resolve: a nearly disconnected pore space).  0 = never. */
    double escalate_k;
} lbm_srv_params;

typedef struct lbm_srv_request {
    int32_t nx, ny, nz;
    const int32_t *solid_ijk; /* [i][j][k], 0 = pore; copied during lbm_srv_submit */
    double nu;                /* kinematic viscosity */
    double g[3];              /* body force */
    int32_t perm_dir;         /* PerDir: 1 = x, 2 = y, 3 = z; the direction whose k is tested */
} lbm_srv_request;

typedef struct lbm_srv_result {
    int64_t ticket;
    double k[3];              /* permeability in lattice units, k_d = Q_d/V nu/g_d (Comput_Perm) */
    double rel_change;        /* last relative change of k[perm_dir-1] */
    int64_t steps;
    int64_t fluid_nodes;
    int32_t converged;        /* 1 converged, 0 stopped at max_steps, -1 non-finite velocity, -2 not preparable */
    int32_t accelerations;    /* Anderson extrapolations applied to this sample */
    int32_t precision;        /* LBM_FP32 or LBM_FP64: the arena the result came from */
    int32_t pad;
    double t_submit, t_admit, t_done;   /* seconds on the server clock */
} lbm_srv_result;

typedef struct lbm_srv_stats {
    int64_t pending, active, done;
    int64_t arena_nodes, arena_used;
    int64_t launches;        /* step launches issued */
    int64_t node_updates;    /* fluid-node updates performed */
    int32_t block;           /* tile size in use */
    int32_t pad;
    double busy_seconds;     /* device time of the step launches (HIP events) */
} lbm_srv_stats;

int lbm_srv_create(const lbm_srv_params *p, lbm_handle **out);
/* queue a request: the solid array is copied and checked here (pore count, arena size, index
 * span); numbering and neighbour deltas are built by a pool of preparation threads */
int lbm_srv_submit(lbm_handle *h, const lbm_srv_request *req, int64_t *ticket);
/* take up to max finished results */
int lbm_srv_poll(lbm_handle *h, lbm_srv_result *out, int32_t max, int32_t *n_out);
/* block until nothing is pending or active (timeout_s <= 0: no limit); LBM_ERR_BUSY on timeout */
int lbm_srv_wait_idle(lbm_handle *h, double timeout_s);
int lbm_srv_get_stats(lbm_handle *h, lbm_srv_stats *out);
/* seconds on the server clock (for latency accounting by callers) */
double lbm_srv_clock(const lbm_handle *h);

/* ------------------------------------------------------------------ launch settings
 * How a handle launches its hot kernels: the workgroup size (nodes per tile), the order in which
 * the tiles of consecutive nodes are dealt to the MI355X's 8 XCDs (each with its own L2), and,
 * for the colour streaming, the variant.  No counterpart in the reference (a CPU program).  By
 * default a handle chooses them itself: at create, on its own geometry, it times every candidate
 * and keeps the fastest (the autotuner; choices are cached per kernel, precision and size class
 * for the life of the process, and boxes under 2^20 fluid nodes keep the defaults).
 * lbm_set_launch fixes a setting, e.g. to compare two; results are identical for every setting. */
#define LBM_KERNEL_SP_STEP 1      /* single-phase step (sp_step_kernel) */
#define LBM_KERNEL_CG_COLLIDE 2   /* colour-gradient collide (cg_collide_kernel) */
#define LBM_KERNEL_CG_COLOUR 3    /* colour streaming (cg_colour_kernel) */
#define LBM_KERNEL_ST_STEP 4      /* passive-solute step (st_step_kernel) */
#define LBM_ORDER_HW 0            /* tiles in the hardware's round-robin order over the XCDs */
#define LBM_ORDER_RUNS8 1         /* each XCD takes runs of 8 consecutive tiles */
#define LBM_ORDER_SLAB 2          /* each XCD sweeps one contiguous slab of tiles */
#define LBM_VARIANT_GATHER 0      /* colour-gradient kernels: 18 neighbour gathers per node through the caches */
#define LBM_VARIANT_STAGED 1      /* the block's neighbourhood staged in LDS first (block 512; the colour
                                     states for the colour streaming, the phase field for the collide) */
typedef struct lbm_launch {
    int32_t block;     /* threads (nodes) per workgroup: 128, 256 or 512 */
    int32_t order;     /* LBM_ORDER_* */
    int32_t variant;   /* LBM_VARIANT_* (colour-gradient kernels; 0 for the others) */
    int32_t tuned;     /* 1: chosen by the autotuner on this geometry, 0: default or set */
    double time_us;    /* the autotuner's time for one launch of the choice (0 if not tuned) */
} lbm_launch;
int lbm_get_launch(const lbm_handle *h, int32_t kernel, lbm_launch *out);
int lbm_set_launch(lbm_handle *h, int32_t kernel, const lbm_launch *in);

/* ------------------------------------------------------------------ common */
const char *lbm_last_error(const lbm_handle *h);   /* NULL handle: last create() failure */
void lbm_destroy(lbm_handle *h);
/* library version string and the device architecture the kernels were built for */
const char *lbm_version(void);
/* Host only, no device use.  *wraps = 1 when some cluster of pore voxels, joined by the 18
 * D3Q19 links with the box periodic on every axis, wraps round the box along axis (0 = x,
 * 1 = y, 2 = z), else 0.  Without such a cluster there is no throughflow along the axis at
 * steady state.  No counterpart in the reference: the server uses it to send samples whose
 * pores do not wrap along the permeability direction to FP64 at admission (their k is the
 * lattice's spurious-current residue, below FP32 resolution); callers can screen samples. */
int lbm_geometry_wraps(int32_t nx, int32_t ny, int32_t nz, const int32_t *solid_ijk, int32_t axis, int32_t *wraps);

#ifdef __cplusplus
}
#endif
#endif
