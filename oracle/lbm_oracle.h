/*
 * lbm_oracle.h -- CPU FP64 restatement of the jianhui-lbm sparse D3Q19 MRT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle.  Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py load it, and only as the
 * checker (or as the timed CPU baseline).  The HIP product never links or calls it.
 *
 * Conventions follow the reference so that outputs compare with its checkpoint files
 * without reordering:
 *   - fluid nodes are numbered 1..count in i-major order (x slowest, z fastest);
 *     index 0 of every per-node array is padding (SP:1172-1193, SP:486-529);
 *   - populations are AoS, f[n*19 + q], in the reference's own velocity ordering
 *     (SP ordering for the single-phase functions, CG ordering for the colour-gradient ones);
 *   - solid_ijk is int32 [i][j][k] (z fastest), 0 = pore, anything else = solid (SP:1058-1065).
 * The reference splits x into MPI slabs joined by a ring; here the box is one rank and x is
 * periodic, which is what that ring gives for the whole box (SP:1954-1999).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* lattice tables in the reference orderings; which = 0 (SP:157-171) or 1 (CG:203-221) */
void orc_lattice(int which, int e[19][3], double w[19], int opp[19]);
/* moment matrix M (SP:56-75 or CG:80-98), built from its basis polynomials */
void orc_moment_matrix(int which, double M[19][19]);
/* inverse of M by Gauss-Jordan elimination (the reference's Comput_MI, SP:932-960) */
int orc_invert19(const double A[19][19], double Ainv[19][19]);

/* Fluid numbering (SP:1172-1193) and SupInv (SP:1261-1277). idx: dense nx*ny*nz map,
 * 0 = solid, else 1-based fluid index; supinv[1..count] = i*ny*nz + j*nz + k.
 * Returns count; supinv may be NULL to only count. */
int orc_number_fluid(int nx, int ny, int nz, const int32_t *solid_ijk, int32_t *idx, int32_t *supinv);

/* ---------------- single phase (SP) ---------------- */
/* relaxation vector S[19] and s_v from the kinematic viscosity (SP:1282-1328) */
void orc_sp_relaxation(double nu, double S[19], double *s_v);
/* f = feq(rho, u0) everywhere, u = u0, rho = rho_init (SP:1343-1376) */
void orc_sp_init(int count, double rho_init, const double u0[3], double *rho, double *u, double *f);
/* collision + push streaming with halfway bounce-back (SP:1494-1926), periodic box */
void orc_sp_collide(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                    const double S[19], const double g[3], const double *rho, const double *u,
                    const double *f, double *F, int nthreads);
/* macroscopic update (SP:2713-2741) */
void orc_sp_macro(int count, const double g[3], const double *F, double *f, double *rho, double *u,
                  double *u0, int nthreads);
/* n steps of collide + macro (the reference time loop body, SP:662-684) */
/* face order in the arrays: x+, x-, y+, y-, z+, z- (the reference's input lines) */
typedef struct orc_bc {
    int sub_bc;
    int pre[6];
    double p[6];
    int vel[6];
    double v[6];
    int no_quirks;   /* 0: the reference's arithmetic; 1: its two quirks corrected (test use) */
} orc_bc;
int orc_sp_boundary(int nx, int ny, int nz, const int32_t *idx, const orc_bc *bc, const double *f,
                    const double *rho, const double *u, double *F);
int orc_sp_run_bc(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                  const double S[19], const double g[3], const orc_bc *bc, double *rho, double *u, double *u0,
                  double *f, double *F, int nsteps, int nthreads);
int orc_sp_run_ex(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                  const double S[19], const double g[3], const orc_bc *bc, const double *pl, double *niu_vn, int step0,
                  double *rho, double *u, double *u0, double *f, double *F, int nsteps, int nthreads);
void orc_sp_run(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                const double S[19], const double g[3], double *rho, double *u, double *u0,
                double *f, double *F, int nsteps, int nthreads);
/* velocity-change error (SP:3625-3704): returns err; *umax and *sum_speed as there */
double orc_sp_error(int count, const double *u, const double *u0, double *umax, double *sum_speed);
/* the same over the reference's ranks: rank r owns nodes [rank_end[r-1], rank_end[r]) */
double orc_sp_error_ranks(int nranks, const int64_t *rank_end, const double *u, const double *u0, double *umax,
                          double *sum_speed);
/* the reference's x-slab split (SP:1085-1108): first x-plane of each rank from the plane pore counts */
void orc_slab_split(int nx, const int32_t *loc_por, int nranks, int32_t *disp);
/* flux sums for the permeability (SP:4617-4645). box = {xn,xp,yn,yp,zn,zp} or NULL */
void orc_flux(int ny, int nz, int count, const int32_t *supinv, const double *u, const int *box,
              double Q[3]);

/* ---------------- colour gradient (CG) ---------------- */
/* viscosity interpolation constants (CG:2368-2377): out = {wl, wg, lg0, l1, l2, g1, g2, delta} */
void orc_cg_visc_consts(double nu_l, double nu_g, double out[8]);
/* init from a phase field (ini_Sat<0 path, CG:2514-2547) with f = M^-1 meq (CG:2644-2681);
 * psi_node[1..count] are the per-node phase values already clamped to [-1, 1] */
void orc_cg_init(int count, const double u0[3], const double *psi_node, double *rho, double *u,
                 double *psi, double *rho_r, double *rho_b, double *rhor, double *rhob, double *f);
typedef struct orc_cg_params {
    double g[3];          /* gxs, gys, gzs */
    double cap_a;         /* CapA, surface tension parameter */
    double psi_solid;     /* ContactAngle_parameter (cos theta) */
    double visc[8];       /* orc_cg_visc_consts output */
    int bodyforce_apply;  /* 0, 1 or -1 (CG:3040) */
    double rel_perm_psi, rel_perm_psi2;
} orc_cg_params;
/* colour-gradient collision, push streaming, recolouring and colour push (CG:2892-3565) */
void orc_cg_collide(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                    const orc_cg_params *p, const double *rho, const double *u, const double *f,
                    double *F, const double *psi, const double *rho_r, const double *rho_b,
                    double *rhor, double *rhob);
/* macroscopic update (CG:3713-3745) */
void orc_cg_macro(int count, const double g[3], const double *F, double *f, double *rho, double *u,
                  double *u0, double *rho_r, double *rho_b, double *rhor, double *rhob, double *psi);
/* phase flux sums (CG:6366-6404) and the saturation count (CG:6640-6678) */
void orc_cg_flux(int ny, int nz, int count, const int32_t *supinv, const double *u, const double *psi,
                 const int *box, double rel_perm_psi, double rel_perm_psi2, double Ql[3], double Qg[3]);
double orc_cg_wet_count(int ny, int nz, int count, const int32_t *supinv, const double *psi, const int *box);

/* ---- passive solute, D3Q7 BGK with rest weight w0 (w0 = 0: the reference's six-velocity
 * scheme, Single_Phase_3D/Solute_Transport_MPI_Software_Spars_6V.cpp).  Arrays 1-based with a
 * padding row 0 like the others: g[(count+1)*7] (index 0 = rest), C[count+1], u[(count+1)*3]. */
void orc_st_init(int count, double w0, const double *C, const double *u, double *g);
void orc_st_step(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv, double w0,
                 double tau_s, const double *u, double *g, double *G, double *C, int nsteps, int nthreads);
double orc_st_mass(int count, const double *C);

#ifdef __cplusplus
}
#endif
#endif
