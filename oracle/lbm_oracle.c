/*
 * lbm_oracle.c -- CPU FP64 restatement of the jianhui-lbm sparse D3Q19 MRT hot path.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle and timed CPU baseline; the HIP product never
 * calls it).  Each function restates one reference routine; the reference line it follows
 * is cited beside it.  SP = Single_Phase_3D/SINGLE_PHASE_MPI_Software_Spars.cpp,
 * CG = MP_3D/Multi_Component_CG_MPI_SPARSE_Software.cpp,
 * 6V.cpp = Single_Phase_3D/Solute_Transport_MPI_Software_Spars_6V.cpp (passive solute).
 *
 * Arithmetic is kept in the reference's evaluation order where that is cheap (left-to-right
 * 19-term sums, the same divisions by c_s^2), so that the oracle tracks the reference to
 * round-off.  It differs from the reference in two documented places only:
 *   - M^-1 is computed here (closed form / elimination) instead of the reference's pasted
 *     hex literals (SP:1834-1870, CG:3364-3400), which are a few ulp off the exact inverse;
 *   - the MPI ring in x (SP:1954-1999) is a periodic wrap, i.e. single-rank semantics.
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, optional OpenMP).
 */
#include "lbm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* D3Q19 velocity sets in the two reference orderings. SP:157-159 orders the diagonals as
 * xy (7-10), yz (11-14), xz (15-18); CG:203-205 orders them as opposite pairs
 * (7,8), (9,10) in xy, 11-14 in xz and 15-18 in yz. */
static const int E_SP[19][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1},
    {0, -1, -1}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1}};
static const int E_CG[19][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1}, {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1}, {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};

void orc_lattice(int which, int e[19][3], double w[19], int opp[19])
{
    const int (*E)[3] = which ? E_CG : E_SP;
    for (int q = 0; q < 19; q++) {
        int e2 = 0;
        for (int a = 0; a < 3; a++) {
            e[q][a] = E[q][a];
            e2 += E[q][a] * E[q][a];
        }
        /* weights 1/3, 1/18, 1/36 written as the reference writes them (SP:161) */
        w[q] = e2 == 0 ? 1.0 / 3.0 : (e2 == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
        for (int r = 0; r < 19; r++)
            if (E[r][0] == -E[q][0] && E[r][1] == -E[q][1] && E[r][2] == -E[q][2]) opp[q] = r;
    }
}

/* Rows of the two moment bases as polynomials of the velocity e (e2 = |e|^2).
 * SP (SP:56-75) is d'Humieres' orthogonal D3Q19 basis; CG (CG:80-98) spans the same
 * polynomial spaces with a non-orthogonal choice in rows 1,2,4,6,8,10,12. */
static double basis_row(int which, int k, const int *e)
{
    double x = e[0], y = e[1], z = e[2];
    double e2 = x * x + y * y + z * z;
    double q9 = 3 * x * x - e2, q11 = y * y - z * z;
    double c1 = which ? e2 - 1 : 19 * e2 - 30;
    double c2 = which ? 3 * e2 * e2 - 6 * e2 + 1 : (21 * e2 * e2 - 53 * e2 + 24) / 2;
    double cv = which ? 3 * e2 - 5 : 5 * e2 - 9;   /* energy-flux factor */
    double ct = which ? 2 * e2 - 3 : 3 * e2 - 5;   /* stress-flux factor */
    switch (k) {
    case 0: return 1;
    case 1: return c1;
    case 2: return c2;
    case 3: return x;
    case 4: return cv * x;
    case 5: return y;
    case 6: return cv * y;
    case 7: return z;
    case 8: return cv * z;
    case 9: return q9;
    case 10: return ct * q9;
    case 11: return q11;
    case 12: return ct * q11;
    case 13: return x * y;
    case 14: return y * z;
    case 15: return x * z;
    case 16: return (y * y - z * z) * x;
    case 17: return (z * z - x * x) * y;
    default: return (x * x - y * y) * z;
    }
}

void orc_moment_matrix(int which, double M[19][19])
{
    const int (*E)[3] = which ? E_CG : E_SP;
    for (int k = 0; k < 19; k++)
        for (int q = 0; q < 19; q++) M[k][q] = basis_row(which, k, E[q]);
}

/* Gauss-Jordan with partial pivoting in long double, rounded once to double. */
int orc_invert19(const double A[19][19], double Ainv[19][19])
{
    long double a[19][38];
    for (int i = 0; i < 19; i++)
        for (int j = 0; j < 38; j++) a[i][j] = j < 19 ? A[i][j] : (j - 19 == i ? 1.0L : 0.0L);
    for (int c = 0; c < 19; c++) {
        int piv = c;
        for (int r = c + 1; r < 19; r++)
            if (fabsl(a[r][c]) > fabsl(a[piv][c])) piv = r;
        if (fabsl(a[piv][c]) < 1e-30L) return -1;
        if (piv != c)
            for (int j = 0; j < 38; j++) {
                long double t = a[c][j];
                a[c][j] = a[piv][j];
                a[piv][j] = t;
            }
        long double d = a[c][c];
        for (int j = 0; j < 38; j++) a[c][j] /= d;
        for (int r = 0; r < 19; r++) {
            if (r == c || a[r][c] == 0.0L) continue;
            long double fct = a[r][c];
            for (int j = 0; j < 38; j++) a[r][j] -= fct * a[c][j];
        }
    }
    for (int i = 0; i < 19; i++)
        for (int j = 0; j < 19; j++) Ainv[i][j] = (double)a[i][j + 19];
    return 0;
}

/* SP: M^-1 = M^T diag(1/|row|^2) because the SP rows are orthogonal. */
static void sp_inverse(double M[19][19], double Mi[19][19])
{
    for (int k = 0; k < 19; k++) {
        double n2 = 0;
        for (int q = 0; q < 19; q++) n2 += M[k][q] * M[k][q];
        for (int q = 0; q < 19; q++) Mi[q][k] = M[k][q] / n2;
    }
}

int orc_number_fluid(int nx, int ny, int nz, const int32_t *solid_ijk, int32_t *idx, int32_t *supinv)
{
    int count = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                long lin = ((long)i * ny + j) * nz + k;
                if (solid_ijk[lin] == 0) {
                    ++count;
                    if (idx) idx[lin] = count;
                    if (supinv) supinv[count] = (int32_t)lin;
                } else if (idx) {
                    idx[lin] = 0;
                }
            }
    return count;
}

/* ===================================================================== SP */

static const double C_S2 = 1.0 / 3.0;   /* lat_c^2/3 with lat_c = 1 (SP:1291-1298) */

void orc_sp_relaxation(double nu, double S[19], double *s_v)
{
    /* tau = nu/(c_s^2 dt) + 1/2, s_v = 1/tau, s_q = 8(2-s_v)/(8-s_v) (SP:1302-1307) */
    double tau = nu / (C_S2 * 1.0) + 0.5;
    double sv = 1 / tau;
    double sq = 8 * (2 - sv) / (8 - sv);
    static const int kind[19] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};
    for (int k = 0; k < 19; k++) S[k] = kind[k] == 0 ? 0.0 : (kind[k] == 1 ? sv : sq);
    if (s_v) *s_v = sv;
}

static double feq_sp(const int *e, double w, double rho, const double *u)
{
    /* w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2), c = 1 (SP:1386-1396) */
    double eu = e[0] * u[0] + e[1] * u[1] + e[2] * u[2];
    double uv = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    return w * rho * (1.0 + 3.0 * eu / 1.0 + 4.5 * eu * eu / 1.0 - 1.5 * uv / 1.0);
}

void orc_sp_init(int count, double rho_init, const double u0[3], double *rho, double *u, double *f)
{
    int e[19][3], opp[19];
    double w[19];
    orc_lattice(0, e, w, opp);
    for (int n = 1; n <= count; n++) {
        for (int a = 0; a < 3; a++) u[n * 3 + a] = u0[a];
        rho[n] = rho_init;
        for (int q = 0; q < 19; q++) f[n * 19 + q] = feq_sp(e[q], w[q], rho[n], u0);
    }
}

/* Guo forcing term for one direction (SP:1519-1593; CG:3048-3120 with g_s) */
static double guo(const int *e, double w, const double *u, const double *g)
{
    double lm0 = ((e[0] - u[0]) * g[0] + (e[1] - u[1]) * g[1] + (e[2] - u[2]) * g[2]) / C_S2;
    double lm1 = (e[0] * u[0] + e[1] * u[1] + e[2] * u[2]) * (e[0] * g[0] + e[1] * g[1] + e[2] * g[2]) /
                 (C_S2 * C_S2);
    return w * (lm0 + lm1);
}

static inline int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

/* pl = {nnf_m, nnf_n} switches on the power-law collision (collision_nnf, SP:2017-2700): from
 * reference step n > 10 on (SP:2397), each node relaxes with its own rates from niu_vn, which it
 * first updates from its strain rate (SP:2399-2491, the "SRT" estimate of SP:2430-2449):
 *   s = 1/(niu_vn/c_s^2 + 1/2);  S_ab = -s/(2 rho c_s^2) sum_q e_a e_b (f_q - feq_q)
 *   niu_vn = m exp((n-1) log sqrt(Sxx^2+Syy^2+Szz^2+2Sxy^2+2Syz^2+2Sxz^2))
 *   s_v = 1/(niu_vn/c_s^2 + 1/2), s_q = 8(2-s_v)/(8-s_v) in place of S (SP:2466-2491). */
static void sp_collide_impl(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                            const double S0[19], const double g[3], const double *rho, const double *u,
                            const double *f, double *F, int nthreads, const double *pl, double *niu_vn, int step)
{
    int e[19][3], opp[19];
    double w[19], M[19][19], Mi[19][19];
    orc_lattice(0, e, w, opp);
    orc_moment_matrix(0, M);
    sp_inverse(M, Mi);
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int ci = 1; ci <= count; ci++) {
        int s = supinv[ci];
        int i = s / (ny * nz), j = (s % (ny * nz)) / nz, m = s % nz;
        const double *fc = f + (long)ci * 19;
        const double *uc = u + (long)ci * 3;
        double gf[19], fe[19], ml[19], S[19];
        for (int q = 0; q < 19; q++) {
            gf[q] = guo(e[q], w[q], uc, g);
            fe[q] = feq_sp(e[q], w[q], rho[ci], uc);
            S[q] = S0[q];
        }
        if (pl && step > 10) {
            const double sv0 = 1.0 / (niu_vn[ci] / (C_S2 * 1.0) + 0.5);
            double sxx = 0, syy = 0, szz = 0, sxy = 0, syz = 0, sxz = 0;
            for (int q = 0; q < 19; q++) {
                const double d = fc[q] - fe[q];
                sxx += e[q][0] * e[q][0] * d;
                syy += e[q][1] * e[q][1] * d;
                szz += e[q][2] * e[q][2] * d;
                sxy += e[q][0] * e[q][1] * d;
                syz += e[q][1] * e[q][2] * d;
                sxz += e[q][0] * e[q][2] * d;
            }
            const double c = -sv0 / (2 * rho[ci] * C_S2 * 1.0);
            sxx *= c;
            syy *= c;
            szz *= c;
            sxy *= c;
            syz *= c;
            sxz *= c;
            niu_vn[ci] = pl[0] * exp((pl[1] - 1) * log(sqrt(sxx * sxx + syy * syy + szz * szz + 2 * sxy * sxy +
                                                            2 * syz * syz + 2 * sxz * sxz)));
            const double sv = 1 / (niu_vn[ci] / (C_S2 * 1.0) + 0.5);
            const double so = 8 * (2 - sv) / (8 - sv);
            static const int kind[19] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};
            for (int k = 0; k < 19; k++) S[k] = kind[k] == 0 ? 0.0 : (kind[k] == 1 ? sv : so);
        }
        /* m = M f, F_hat = (I - S/2) M GuoF, meq = M feq, relax (SP:1642-1829) */
        for (int k = 0; k < 19; k++) {
            double mk = M[k][0] * fc[0], fk = M[k][0] * gf[0], ek = M[k][0] * fe[0];
            for (int q = 1; q < 19; q++) {
                mk += M[k][q] * fc[q];
                fk += M[k][q] * gf[q];
                ek += M[k][q] * fe[q];
            }
            fk *= (1 - 0.5 * S[k]);
            ml[k] = mk - S[k] * (mk - ek) + 1.0 * fk;
        }
        /* f* = M^-1 m* and push to the downstream node, or bounce back (SP:1834-1926) */
        for (int q = 0; q < 19; q++) {
            double v = Mi[q][0] * ml[0];
            for (int k = 1; k < 19; k++) v += Mi[q][k] * ml[k];
            int ip = wrap(i + e[q][0], nx), jp = wrap(j + e[q][1], ny), kp = wrap(m + e[q][2], nz);
            int t = idx[((long)ip * ny + jp) * nz + kp];
            if (t > 0)
                F[(long)t * 19 + q] = v;
            else
                F[(long)ci * 19 + opp[q]] = v;
        }
    }
}

void orc_sp_collide(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                    const double S[19], const double g[3], const double *rho, const double *u,
                    const double *f, double *F, int nthreads)
{
    sp_collide_impl(nx, ny, nz, count, idx, supinv, S, g, rho, u, f, F, nthreads, NULL, NULL, 0);
}

void orc_sp_macro(int count, const double g[3], const double *F, double *f, double *rho, double *u,
                  double *u0, int nthreads)
{
    int e[19][3], opp[19];
    double w[19];
    orc_lattice(0, e, w, opp);
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int n = 1; n <= count; n++) {
        double r = 0, ux = 0, uy = 0, uz = 0;
        for (int a = 0; a < 3; a++) u0[n * 3 + a] = u[n * 3 + a];
        for (int q = 0; q < 19; q++) {
            double v = F[(long)n * 19 + q];
            f[(long)n * 19 + q] = v;
            r += v;
            ux += e[q][0] * v;
            uy += e[q][1] * v;
            uz += e[q][2] * v;
        }
        rho[n] = r;
        u[n * 3 + 0] = (ux + 1.0 * g[0] / 2) / r;
        u[n * 3 + 1] = (uy + 1.0 * g[1] / 2) / r;
        u[n * 3 + 2] = (uz + 1.0 * g[2] / 2) / r;
    }
}

void orc_sp_run(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                const double S[19], const double g[3], double *rho, double *u, double *u0,
                double *f, double *F, int nsteps, int nthreads)
{
    for (int s = 0; s < nsteps; s++) {
        orc_sp_collide(nx, ny, nz, count, idx, supinv, S, g, rho, u, f, F, nthreads);
        orc_sp_macro(count, g, F, f, rho, u, u0, nthreads);
    }
}

/* ------------------------------------------------------------------ SP boundaries */
/* Pressure and velocity boundaries on the faces of the box, applied to the streamed F between
 * collision and comput_macro_variables (SP:676-682): boundary_pressure (SP:3076-3620), then
 * boundary_velocity (SP:2751-3075).  Faces are visited in the reference's order y+, y-, z+,
 * z-, x+, x-, so the last face that owns a node decides its state.  The inward neighbour nb of
 * a face node is the next node along the face normal (j = NY-1 for y+, i = 1 for x-, ...), and
 * f, rho, u at nb are the pre-collision values of this step.  Single-rank semantics: the x
 * faces are i = 0 and i = nx-1 of the whole box.  The reference's quirks are kept:
 *   - x- face with Sub_BC 1 and 3: u_ls[2] takes u[nb][1] (SP:3254, 3613);
 *   - x faces with Sub_BC 2 and a solid nb: the equilibrium is written to the node at k = NZ of
 *     the same (i, j) column instead of the face node (SP:3451, 3475), which keeps its streamed
 *     populations; the write to the column's own k = NZ node is overwritten when the loop
 *     reaches it.
 * (no_quirks = 1 replaces both by the evident intent; tests use it to show that the reference's
 * outputs tell the two apart.)
 * Velocity Sub_BC 2 and 3 are not restated (their x+ face indexes Solid[nx_l-1][k][j],
 * SP:2953, 3051); the function returns -1 for them. */
static double feq_u(int q, const int e[19][3], const double w[19], double rho, double ux, double uy, double uz)
{
    const double u[3] = {ux, uy, uz};
    return feq_sp(e[q], w[q], rho, u);
}

int orc_sp_boundary(int nx, int ny, int nz, const int32_t *idx, const orc_bc *bc, const double *f,
                    const double *rho, const double *u, double *F)
{
    int e[19][3], opp[19];
    double w[19], M[19][19], Mi[19][19];
    orc_lattice(0, e, w, opp);
    orc_moment_matrix(0, M);
    sp_inverse(M, Mi);
    const int NZ = nz - 1;
#define ND(i, j, k) idx[((long)(i) * ny + (j)) * nz + (k)]
    int any_p = 0, any_v = 0;
    for (int a = 0; a < 6; a++) {
        any_p |= bc->pre[a] == 1;
        any_v |= bc->vel[a] == 1;
    }
    if (any_v && (bc->sub_bc == 2 || bc->sub_bc == 3)) return -1;
    /* faces in visiting order: code, normal axis, face plane, inward neighbour plane */
    static const int order[6] = {2, 3, 4, 5, 0, 1};   /* yp yn zp zn xp xn in xp,xn,yp,yn,zp,zn numbering */
    if (any_p) {
        for (int oi = 0; oi < 6; oi++) {
            const int fc = order[oi];
            if (bc->pre[fc] != 1) continue;
            const double rb = bc->p[fc];
            const int ax = fc / 2;                   /* 0 x, 1 y, 2 z */
            const int dim[3] = {nx, ny, nz};
            const int plane = (fc % 2 == 0) ? dim[ax] - 1 : 0;
            const int inward = (fc % 2 == 0) ? plane - 1 : 1;
            const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
            for (int p1 = 0; p1 < dim[a1]; p1++)
                for (int p2 = 0; p2 < dim[a2]; p2++) {
                    int c[3], cn[3];
                    c[ax] = plane;
                    c[a1] = p1;
                    c[a2] = p2;
                    cn[ax] = inward;
                    cn[a1] = p1;
                    cn[a2] = p2;
                    const int b = ND(c[0], c[1], c[2]);
                    if (b <= 0) continue;
                    const int nb = ND(cn[0], cn[1], cn[2]);
                    double *Fb = F + (long)b * 19;
                    if (bc->sub_bc == 0 || nb <= 0) {
                        if (bc->sub_bc == 2 && nb <= 0 && ax == 0 && !bc->no_quirks) {
                            const int b2 = ND(c[0], c[1], NZ);   /* SP:3451, 3475 */
                            for (int q = 0; q < 19; q++) F[(long)b2 * 19 + q] = feq_u(q, e, w, rb, 0, 0, 0);
                            continue;
                        }
                        for (int q = 0; q < 19; q++) Fb[q] = feq_u(q, e, w, rb, 0, 0, 0);
                        continue;
                    }
                    const double *un = u + (long)nb * 3;
                    const double ulz = (fc == 1 && !bc->no_quirks) ? un[1] : un[2];   /* x- face: u_ls[2] = u[nb][1] */
                    if (bc->sub_bc == 1) {
                        for (int q = 0; q < 19; q++) Fb[q] = feq_u(q, e, w, rb, un[0], un[1], ulz);
                    } else if (bc->sub_bc == 2) {
                        const double *fn = f + (long)nb * 19;
                        double ml[19];
                        for (int k = 0; k < 19; k++) {
                            ml[k] = 0;
                            for (int q = 0; q < 19; q++) ml[k] += M[k][q] * fn[q];
                        }
                        ml[0] = rb;
                        for (int q = 0; q < 19; q++) {
                            Fb[q] = 0;
                            for (int k = 0; k < 19; k++) Fb[q] += Mi[q][k] * ml[k];
                        }
                    } else {
                        const double *fn = f + (long)nb * 19;
                        for (int q = 0; q < 19; q++)
                            Fb[q] = feq_u(q, e, w, rb, un[0], un[1], ulz) + fn[q] -
                                    feq_u(q, e, w, rho[nb], un[0], un[1], un[2]);
                    }
                }
        }
    }
    if (any_v) {
        for (int oi = 0; oi < 6; oi++) {
            const int fc = order[oi];
            if (bc->vel[fc] != 1) continue;
            const int ax = fc / 2;
            double uf[3] = {0, 0, 0};
            uf[ax] = bc->v[fc];
            const int dim[3] = {nx, ny, nz};
            const int plane = (fc % 2 == 0) ? dim[ax] - 1 : 0;
            const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
            for (int p1 = 0; p1 < dim[a1]; p1++)
                for (int p2 = 0; p2 < dim[a2]; p2++) {
                    int c[3];
                    c[ax] = plane;
                    c[a1] = p1;
                    c[a2] = p2;
                    const int b = ND(c[0], c[1], c[2]);
                    if (b <= 0) continue;   /* Sub_BC 0 writes the padding record for solid x-face sites */
                    for (int q = 0; q < 19; q++) F[(long)b * 19 + q] = feq_u(q, e, w, 1.0, uf[0], uf[1], uf[2]);
                }
        }
    }
#undef ND
    return 0;
}

int orc_sp_run_bc(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                  const double S[19], const double g[3], const orc_bc *bc, double *rho, double *u, double *u0,
                  double *f, double *F, int nsteps, int nthreads)
{
    for (int s = 0; s < nsteps; s++) {
        orc_sp_collide(nx, ny, nz, count, idx, supinv, S, g, rho, u, f, F, nthreads);
        if (orc_sp_boundary(nx, ny, nz, idx, bc, f, rho, u, F) != 0) return -1;
        orc_sp_macro(count, g, F, f, rho, u, u0, nthreads);
    }
    return 0;
}

/* The time loop of SP:662-684 with either option: the power-law collision (pl, niu_vn; step0
 * is the reference's loop counter n of the first step) and the faces (bc).  Either may be NULL. */
int orc_sp_run_ex(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                  const double S[19], const double g[3], const orc_bc *bc, const double *pl, double *niu_vn, int step0,
                  double *rho, double *u, double *u0, double *f, double *F, int nsteps, int nthreads)
{
    for (int s = 0; s < nsteps; s++) {
        sp_collide_impl(nx, ny, nz, count, idx, supinv, S, g, rho, u, f, F, nthreads, pl, niu_vn, step0 + s);
        if (bc && orc_sp_boundary(nx, ny, nz, idx, bc, f, rho, u, F) != 0) return -1;
        orc_sp_macro(count, g, F, f, rho, u, u0, nthreads);
    }
    return 0;
}

/* Error over the reference's ranks (SP:3625-3704): each rank loops its own nodes i = 1..Count-1
 * and so skips its own last node (SP:3651); it forms its own err_r = sqrt(t1)/(sqrt(t2)+1e-30)
 * (SP:3666-3669); rank 0 takes the largest err_r and the largest per-rank u_max and sums the
 * per-rank speed sums (SP:3677-3685).  rank_end[r] is the exclusive end of rank r's node range
 * in the global 0-based node order (ranks own consecutive x-slabs, so consecutive node ids);
 * u and u0 carry the padding row 0, node n at row n + 1. */
double orc_sp_error_ranks(int nranks, const int64_t *rank_end, const double *u, const double *u0, double *umax,
                          double *sum_speed)
{
    double err = 0, um_all = 0, t3_all = 0;
    int64_t lo = 0;
    for (int r = 0; r < nranks; r++) {
        const int64_t hi = rank_end[r];
        double t1 = 0, t2 = 0, t3 = 0, um = 0;
        for (int64_t n = lo + 1; n < hi; n++) {   /* rows lo+1 .. hi-1: the rank's nodes but its last */
            const double *a = u + n * 3, *b = u0 + n * 3;
            t1 += (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
            const double s2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
            t2 += s2;
            t3 += sqrt(s2);
            if (s2 > um) um = s2;
        }
        const double e = sqrt(t1) / (sqrt(t2) + 1e-30);
        um = sqrt(um);
        if (r == 0 || e > err) err = e;
        if (um > um_all) um_all = um;
        t3_all += t3;
        lo = hi;
    }
    if (umax) *umax = um_all;
    if (sum_speed) *sum_speed = t3_all;
    return err;
}

double orc_sp_error(int count, const double *u, const double *u0, double *umax, double *sum_speed)
{
    /* one rank: the loop stops at count-1, the last node is skipped (SP:3651) */
    const int64_t end = count;
    return orc_sp_error_ranks(1, &end, u, u0, umax, sum_speed);
}

/* The reference's x-slab split over ranks (Parallelize_Geometry, SP:1085-1108): with
 * loc_por[i] the pore count of x-plane i and ave = total/nranks, the cut i+1 is placed where the
 * running plane sum first reaches ave*(i+1), before or after the plane that crosses it,
 * whichever sum is closer; the last rank takes the remaining planes.  disp[r] = first plane of
 * rank r (nranks entries). */
void orc_slab_split(int nx, const int32_t *loc_por, int nranks, int32_t *disp)
{
    int64_t total = 0;
    for (int i = 0; i < nx; i++) total += loc_por[i];
    const double ave = (double)total / nranks;
    disp[0] = 0;
    for (int i = 0; i < nranks - 1; i++) {
        int64_t pre = 0, aft = 0;
        int n_i = 0;
        while (aft < ave * (i + 1)) {
            pre = aft;
            aft += loc_por[n_i];
            n_i += 1;
        }
        disp[i + 1] = ((double)(aft - ave * (i + 1)) > (double)(ave * (i + 1) - pre)) ? n_i - 1 : n_i;
    }
}

static int in_box(int s, int ny, int nz, const int *box)
{
    if (!box) return 1;
    int i = s / (ny * nz), j = (s % (ny * nz)) / nz, k = s % nz;
    return i >= box[0] && i <= box[1] && j >= box[2] && j <= box[3] && k >= box[4] && k <= box[5];
}

void orc_flux(int ny, int nz, int count, const int32_t *supinv, const double *u, const int *box,
              double Q[3])
{
    Q[0] = Q[1] = Q[2] = 0;
    for (int n = 1; n <= count; n++) {
        if (!in_box(supinv[n], ny, nz, box)) continue;
        Q[0] += u[n * 3 + 0];
        Q[1] += u[n * 3 + 1];
        Q[2] += u[n * 3 + 2];
    }
}

/* ===================================================================== CG */

void orc_cg_visc_consts(double nu_l, double nu_g, double out[8])
{
    /* CG:2368-2377, delta = 0.1 */
    double delta = 0.1;
    double wl = 1.0 / (nu_l / (C_S2 * 1.0) + 0.5);
    double wg = 1.0 / (nu_g / (C_S2 * 1.0) + 0.5);
    double lg0 = 2 * wl * wg / (wl + wg);
    double l1 = 2 * (wl - lg0) / delta;
    double l2 = -l1 / (2 * delta);
    double g1 = 2 * (lg0 - wg) / delta;
    double g2 = g1 / (2 * delta);
    out[0] = wl; out[1] = wg; out[2] = lg0; out[3] = l1; out[4] = l2; out[5] = g1; out[6] = g2;
    out[7] = delta;
}

/* meq in the CG basis with rho_0 = 1, plus the surface-tension terms (CG:2664-2673, 3137-3143) */
static void cg_meq(double rho, const double *u, double capa_cc, const double *nv, double meq[19])
{
    double ux = u[0], uy = u[1], uz = u[2], nx = nv[0], ny = nv[1], nz = nv[2];
    for (int k = 0; k < 19; k++) meq[k] = 0;
    meq[0] = rho;
    meq[3] = 1.0 * ux;
    meq[5] = 1.0 * uy;
    meq[7] = 1.0 * uz;
    meq[1] = 1.0 * (ux * ux + uy * uy + uz * uz) + capa_cc;
    meq[9] = 1.0 * (2 * ux * ux - uy * uy - uz * uz) + 0.5 * capa_cc * (2 * nx * nx - ny * ny - nz * nz);
    meq[11] = 1.0 * (uy * uy - uz * uz) + 0.5 * capa_cc * (ny * ny - nz * nz);
    meq[13] = 1.0 * ux * uy + 0.5 * capa_cc * (nx * ny);
    meq[14] = 1.0 * uy * uz + 0.5 * capa_cc * (ny * nz);
    meq[15] = 1.0 * ux * uz + 0.5 * capa_cc * (nx * nz);
}

void orc_cg_init(int count, const double u0[3], const double *psi_node, double *rho, double *u,
                 double *psi, double *rho_r, double *rho_b, double *rhor, double *rhob, double *f)
{
    double M[19][19], Mi[19][19], meq[19];
    const double zero[3] = {0, 0, 0};
    orc_moment_matrix(1, M);
    orc_invert19(M, Mi);
    for (int n = 1; n <= count; n++) {
        for (int a = 0; a < 3; a++) u[n * 3 + a] = u0[a];
        psi[n] = psi_node[n];
        rho[n] = 1.0;
        rho_r[n] = (psi[n] * rho[n] + rho[n]) / 2;
        rho_b[n] = rho[n] - rho_r[n];
        rhor[n] = 0;
        rhob[n] = 0;
        cg_meq(rho[n], u0, 0.0, zero, meq);
        for (int q = 0; q < 19; q++) {
            double v = 0;
            for (int k = 0; k < 19; k++) v += Mi[q][k] * meq[k];
            f[(long)n * 19 + q] = v;
        }
    }
}

void orc_cg_collide(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv,
                    const orc_cg_params *p, const double *rho, const double *u, const double *f,
                    double *F, const double *psi, const double *rho_r, const double *rho_b,
                    double *rhor, double *rhob)
{
    int e[19][3], opp[19];
    double w[19], M[19][19], Mi[19][19];
    orc_lattice(1, e, w, opp);
    orc_moment_matrix(1, M);
    orc_invert19(M, Mi);
    const double wl = p->visc[0], wg = p->visc[1], lg0 = p->visc[2], l1 = p->visc[3],
                 l2 = p->visc[4], g1 = p->visc[5], g2 = p->visc[6], delta = p->visc[7];
    for (int ci = 1; ci <= count; ci++) {
        int s = supinv[ci];
        int i = s / (ny * nz), j = (s % (ny * nz)) / nz, m = s % nz;
        const double *fc = f + (long)ci * 19;
        const double *uc = u + (long)ci * 3;
        /* colour gradient C = sum 3 w e psi(x+e), psi_solid at solid sites (CG:2903-3012) */
        double C[3] = {0, 0, 0};
        int ind_s = 0;
        for (int q = 0; q < 19; q++) {
            int ip = wrap(i + e[q][0], nx), jp = wrap(j + e[q][1], ny), kp = wrap(m + e[q][2], nz);
            int t = idx[((long)ip * ny + jp) * nz + kp];
            double ps;
            if (t > 0)
                ps = psi[t];
            else {
                ind_s = 1;
                ps = p->psi_solid;
            }
            for (int a = 0; a < 3; a++) C[a] += 3.0 * w[q] * e[q][a] * ps;
        }
        double uu = uc[0] * uc[0] + uc[1] * uc[1] + uc[2] * uc[2];
        /* wall suppression (CG:3014-3015) */
        if (sqrt((rho_r[ci] - rho_b[ci]) * (rho_r[ci] - rho_b[ci])) >= 0.9 && ind_s) C[0] = C[1] = C[2] = 0;
        double cc = sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
        double nv[3] = {0, 0, 0};
        if (cc > 0) {
            nv[0] = C[0] / cc;
            nv[1] = C[1] / cc;
            nv[2] = C[2] / cc;
        }
        /* Guo forcing, gated per phase when bodyforce_apply != 0 (CG:3040-3120) */
        double gf[19];
        int gate = (p->bodyforce_apply == 1 && psi[ci] < p->rel_perm_psi) ||
                   (p->bodyforce_apply == -1 && psi[ci] > -p->rel_perm_psi2);
        for (int q = 0; q < 19; q++) gf[q] = gate ? 0.0 : guo(e[q], w[q], uc, p->g);
        double meq[19];
        cg_meq(rho[ci], uc, p->cap_a * cc, nv, meq);
        /* viscosity interpolation in psi (CG:3163-3176) */
        double ps = psi[ci], sv;
        if (ps > 0)
            sv = ps > delta ? wl : lg0 + l1 * ps + l2 * ps * ps;
        else
            sv = ps < -delta ? wg : lg0 + g1 * ps + g2 * ps * ps;
        double so = 8 * (2 - sv) / (8 - sv);
        static const int kind[19] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};
        double S[19];
        for (int k = 0; k < 19; k++) S[k] = kind[k] == 0 ? 0.0 : (kind[k] == 1 ? sv : so);
        double ml[19];
        for (int k = 0; k < 19; k++) {
            double mk = M[k][0] * fc[0], fk = M[k][0] * gf[0];
            for (int q = 1; q < 19; q++) {
                mk += M[k][q] * fc[q];
                fk += M[k][q] * gf[q];
            }
            fk *= (1 - 0.5 * S[k]);
            ml[k] = mk - S[k] * (mk - meq[k]) + 1.0 * fk;
        }
        double gr[19], gb[19];
        for (int q = 0; q < 19; q++) {
            double v = Mi[q][0] * ml[0];
            for (int k = 1; k < 19; k++) v += Mi[q][k] * ml[k];
            int ip = wrap(i + e[q][0], nx), jp = wrap(j + e[q][1], ny), kp = wrap(m + e[q][2], nz);
            int t = idx[((long)ip * ny + jp) * nz + kp];
            if (t > 0)
                F[(long)t * 19 + q] = v;
            else
                F[(long)ci * 19 + opp[q]] = v;
            /* colour distributions with the pre-collision u (CG:3477-3485) */
            double eu = e[q][0] * uc[0] + e[q][1] * uc[1] + e[q][2] * uc[2];
            double phi = 1 + 3 * eu / 1.0 + 4.5 * eu * eu / 1.0 - 1.5 * uu / 1.0;
            gr[q] = w[q] * rho_r[ci] * phi;
            gb[q] = w[q] * rho_b[ci] * phi;
        }
        /* recolouring over opposite pairs (kk, kk+1) (CG:3493-3508) */
        if (cc > 0)
            for (int kk = 1; kk < 19; kk += 2) {
                double ef = e[kk][0] * C[0] + e[kk][1] * C[1] + e[kk][2] * C[2];
                double d = gr[kk] < gr[kk + 1] ? gr[kk] : gr[kk + 1];
                d = d < gb[kk] ? d : gb[kk];
                d = d < gb[kk + 1] ? d : gb[kk + 1];
                d *= ef / cc;
                gr[kk] += d;
                gr[kk + 1] -= d;
                gb[kk] -= d;
                gb[kk + 1] += d;
            }
        /* colour push-add, own site when the target is solid (CG:3513-3565) */
        for (int q = 0; q < 19; q++) {
            int ip = wrap(i + e[q][0], nx), jp = wrap(j + e[q][1], ny), kp = wrap(m + e[q][2], nz);
            int t = idx[((long)ip * ny + jp) * nz + kp];
            int dst = t > 0 ? t : ci;
            rhor[dst] += gr[q];
            rhob[dst] += gb[q];
        }
    }
}

void orc_cg_macro(int count, const double g[3], const double *F, double *f, double *rho, double *u,
                  double *u0, double *rho_r, double *rho_b, double *rhor, double *rhob, double *psi)
{
    int e[19][3], opp[19];
    double w[19];
    orc_lattice(1, e, w, opp);
    for (int n = 1; n <= count; n++) {
        double r = 0, ux = 0, uy = 0, uz = 0;
        for (int a = 0; a < 3; a++) u0[n * 3 + a] = u[n * 3 + a];
        for (int q = 0; q < 19; q++) {
            double v = F[(long)n * 19 + q];
            f[(long)n * 19 + q] = v;
            r += v;
            ux += e[q][0] * v;
            uy += e[q][1] * v;
            uz += e[q][2] * v;
        }
        rho[n] = r;
        rho_r[n] = rhor[n];
        rho_b[n] = rhob[n];
        rhor[n] = 0;
        rhob[n] = 0;
        u[n * 3 + 0] = (ux + 1.0 * g[0] / 2) / r;
        u[n * 3 + 1] = (uy + 1.0 * g[1] / 2) / r;
        u[n * 3 + 2] = (uz + 1.0 * g[2] / 2) / r;
        psi[n] = (rho_r[n] - rho_b[n]) / (rho_r[n] + rho_b[n]);
    }
}

void orc_cg_flux(int ny, int nz, int count, const int32_t *supinv, const double *u, const double *psi,
                 const int *box, double rel_perm_psi, double rel_perm_psi2, double Ql[3], double Qg[3])
{
    for (int a = 0; a < 3; a++) Ql[a] = Qg[a] = 0;
    for (int n = 1; n <= count; n++) {
        if (!in_box(supinv[n], ny, nz, box)) continue;
        if (psi[n] >= rel_perm_psi)
            for (int a = 0; a < 3; a++) Ql[a] += u[n * 3 + a];
        if (psi[n] <= -rel_perm_psi2)
            for (int a = 0; a < 3; a++) Qg[a] += u[n * 3 + a];
    }
}

double orc_cg_wet_count(int ny, int nz, int count, const int32_t *supinv, const double *psi, const int *box)
{
    double sl = 0;
    for (int n = 1; n <= count; n++)
        if (in_box(supinv[n], ny, nz, box) && psi[n] >= 0) sl += 1;
    return sl;
}

/* ------------------------------------------------------------------ passive solute */
/* D3Q7 lattice in the reference's six-velocity order (e2, 6V.cpp:130) plus the rest direction:
 * 0 rest, 1 +x, 2 -x, 3 +y, 4 -y, 5 +z, 6 -z; opposite pairs LR (6V.cpp:138). */
static const int ST_E[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
static const int ST_LR[7] = {0, 2, 1, 4, 3, 6, 5};

/* g_eq_q = w_q C (1 + e_q.u / c_s^2), w_q = (1 - w0)/6, c_s^2 = (1 - w0)/3, w_rest = w0.  With
 * w0 = 0 this is 1/6 C (1 + 3 e.u) of init and collision (6V.cpp:1306-1310, 1473). */
static double st_feq(int q, double w0, double C, const double *u)
{
    if (q == 0) return w0 * C;
    const double w = (1.0 - w0) / 6.0, cs2 = (1.0 - w0) / 3.0;
    const double eu = ST_E[q][0] * u[0] + ST_E[q][1] * u[1] + ST_E[q][2] * u[2];
    return w * C * (1.0 + eu / cs2);
}

void orc_st_init(int count, double w0, const double *C, const double *u, double *g)
{
    for (int n = 1; n <= count; n++)
        for (int q = 0; q < 7; q++) g[(long)n * 7 + q] = st_feq(q, w0, C[n], u + (long)n * 3);
}

/* nsteps x (collision + streaming + macro): BGK relaxation g* = g - (g - g_eq)/tau_s
 * (6V.cpp:1466-1474), push to the downstream node or bounce back into the opposite population
 * of the node itself when it is solid (6V.cpp:1481-1535), then C = sum of the populations
 * (6V.cpp:1625-1640).  The rest population stays on the node.  C is updated in place.
 * Every slot of G receives exactly one value (from its upstream node or by bounce-back from
 * the node itself), so the node loop runs in parallel without changing the result. */
void orc_st_step(int nx, int ny, int nz, int count, const int32_t *idx, const int32_t *supinv, double w0,
                 double tau_s, const double *u, double *g, double *G, double *C, int nsteps, int nthreads)
{
    for (int step = 0; step < nsteps; step++) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
        for (int ci = 1; ci <= count; ci++) {
            int s = supinv[ci];
            int i = s / (ny * nz), j = (s % (ny * nz)) / nz, m = s % nz;
            for (int q = 0; q < 7; q++) {
                const double gq = g[(long)ci * 7 + q];
                const double v = gq - (gq - st_feq(q, w0, C[ci], u + (long)ci * 3)) / tau_s;
                if (q == 0) {
                    G[(long)ci * 7] = v;
                    continue;
                }
                int ip = wrap(i + ST_E[q][0], nx), jp = wrap(j + ST_E[q][1], ny), kp = wrap(m + ST_E[q][2], nz);
                int t = idx[((long)ip * ny + jp) * nz + kp];
                if (t > 0)
                    G[(long)t * 7 + q] = v;
                else
                    G[(long)ci * 7 + ST_LR[q]] = v;
            }
        }
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
        for (int n = 1; n <= count; n++) {
            double c = 0;
            for (int q = 0; q < 7; q++) {
                g[(long)n * 7 + q] = G[(long)n * 7 + q];
                c += G[(long)n * 7 + q];
            }
            C[n] = c;
        }
    }
}

double orc_st_mass(int count, const double *C)
{
    double s = 0;
    for (int n = 1; n <= count; n++) s += C[n];
    return s;
}
