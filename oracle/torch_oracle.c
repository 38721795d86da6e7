/*
 * torch_oracle.c -- CPU FP64 restatement of TORCH's hydro + thermodynamics
 * step.  TEST INFRASTRUCTURE ONLY (see torch_oracle.h): the HIP product path
 * never links or calls this file.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (no -march, no -ffast-math), so
 * that every + - * / is an IEEE double operation in source order, as in the
 * reference's g++ x86-64 build.
 *
 * Each function names the reference file:line it restates.  Expressions keep
 * the reference's operand order and parenthesisation; where the reference
 * uses Eigen 3.2.1 fixed-size algebra, the order Eigen's unrollers produce is
 * written out explicitly:
 *   cross  (a1b2-a2b1, a2b0-a0b2, a0b1-a1b0)   lib/Eigen/src/Geometry/OrthoMethods.h:34-38
 *   dot / squaredNorm   x0 + (x1 + x2)          lib/Eigen/src/Core/Redux.h:77-106
 *   mat*vec, mat*mat    (l0r0 + l1r1) + l2r2     lib/Eigen/src/Core/products/CoeffBasedProduct.h:239-257
 *   normalize           x_i * (1 / norm)        lib/Eigen/src/Core/Dot.h:154-157 (operator/= by a
 *                       scalar multiplies by its reciprocal in Eigen 3.2)
 */
#include "torch_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DEN ORC_DEN
#define PRE ORC_PRE
#define HII ORC_HII
#define ADV ORC_ADV
#define VEL ORC_VEL
#define NV ORC_NV

/* std::min / std::max semantics (NaN handling follows the comparison). */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* Slope limiters: reference src/Integrators/SlopeLimiter.cpp:7-54            */
/* ------------------------------------------------------------------------ */
static inline double minmod(double a, double b) {
    if (a * b > 0) return fabs(a) < fabs(b) ? a : b;
    return 0;
}
static inline double maxmod(double a, double b) {
    if (a * b > 0) return fabs(a) > fabs(b) ? a : b;
    return 0;
}

double orc_limiter(int limiter, double a, double b) {
    switch (limiter) {
    case ORC_LIM_MC:       return minmod((a + b) / 2.0, minmod(2.0 * a, 2.0 * b));
    case ORC_LIM_SUPERBEE: return maxmod(minmod(b, 2.0 * a), minmod(2.0 * b, a));
    case ORC_LIM_MINMOD:   return minmod(a, b);
    case ORC_LIM_MAXMOD:   return maxmod(a, b);
    case ORC_LIM_LEER:
        if (b == 0) return 0;
        else { double r = fabs(a / b); return (a + b * r) / (1.0 + r); }
    case ORC_LIM_OSPRE:
        if (a * b <= 1.0e-30) return 0;
        return 1.5 * b * (a * a + a * b) / (a * a + a * b + b * b);
    default: /* ORC_LIM_ALBADA, SlopeLimiter.cpp:52-54 */
        return a * b <= 1.0e-30 ? 0 : a * b * (a + b) / (a * a + b * b);
    }
}

/* ------------------------------------------------------------------------ */
/* Conversions: reference src/Fluid/GridCell.cpp:163-213 (nd = 3)           */
/* ------------------------------------------------------------------------ */
void orc_UfromQ(double u[7], const double q[7], double gamma) {          /* :163-173 */
    double ke = 0;
    u[DEN] = q[DEN];
    for (int dim = 0; dim < 3; ++dim) {
        u[VEL + dim] = q[VEL + dim] * q[DEN];
        ke += 0.5 * q[DEN] * q[VEL + dim] * q[VEL + dim];
    }
    u[PRE] = q[PRE] / (gamma - 1.0) + ke;
    u[HII] = q[HII] * q[DEN];
    u[ADV] = q[ADV] * q[DEN];
}

void orc_QfromU(double q[7], const double u[7], double gamma) {          /* :175-185 */
    q[DEN] = u[DEN];
    double ke = 0;
    for (int dim = 0; dim < 3; ++dim) {
        q[VEL + dim] = u[VEL + dim] / u[DEN];
        ke += 0.5 * u[VEL + dim] * u[VEL + dim] / u[DEN];
    }
    q[PRE] = (u[PRE] - ke) * (gamma - 1.0);
    q[HII] = u[HII] / u[DEN];
    q[ADV] = u[ADV] / u[DEN];
}

void orc_FfromQ(double f[7], const double q[7], double gamma, int dim) { /* :200-213 */
    f[DEN] = q[DEN] * q[VEL + dim];
    double ke = 0;
    for (int id = 0; id < 3; ++id) {
        ke += q[VEL + id] * q[VEL + id];
        f[VEL + (dim + id) % 3] = q[DEN] * q[VEL + (dim + id) % 3] * q[VEL + dim];
    }
    ke *= 0.5 * q[DEN];
    f[VEL + dim] += q[PRE];
    double g2 = gamma / (gamma - 1.0);
    f[PRE] = q[VEL + dim] * (g2 * q[PRE] + ke);
    f[HII] = q[DEN] * q[VEL + dim] * q[HII];
    f[ADV] = q[DEN] * q[VEL + dim] * q[ADV];
}

/* ------------------------------------------------------------------------ */
/* Riemann solvers: reference src/Integrators/Riemann.cpp                    */
/* ------------------------------------------------------------------------ */
static void wave_speeds(double a_l2, double a_r2, const double* Ql, const double* Qr, int dim,
                        double* S_l, double* S_r) {                        /* :27-42 */
    double sqrtrho_l = sqrt(Ql[DEN]);
    double sqrtrho_r = sqrt(Qr[DEN]);
    double u_tilde = (sqrtrho_l * Ql[VEL + dim] + sqrtrho_r * Qr[VEL + dim]) / (sqrtrho_l + sqrtrho_r);
    double nu2 = 0.5 * sqrtrho_r * sqrtrho_l / ((sqrtrho_l + sqrtrho_r) * (sqrtrho_l + sqrtrho_r));
    double dsqrd = (sqrtrho_l * a_l2 + sqrtrho_r * a_r2) / (sqrtrho_l + sqrtrho_r);
    dsqrd += nu2 * (Qr[VEL + dim] - Ql[VEL + dim]) * (Qr[VEL + dim] - Ql[VEL + dim]);
    double d = sqrt(dsqrd);
    *S_l = u_tilde - d;
    *S_r = u_tilde + d;
}

static int hllc(double F[7], const double* Ql, const double* Qr, double a_l2, double a_r2,
                double gamma, int dim) {                                    /* :52-104 */
    double S_l, S_r;
    wave_speeds(a_l2, a_r2, Ql, Qr, dim, &S_l, &S_r);
    if (S_l >= 0) {
        orc_FfromQ(F, Ql, gamma, dim);
    } else if (S_r <= 0) {
        orc_FfromQ(F, Qr, gamma, dim);
    } else {
        double S_c = Qr[PRE] - Ql[PRE];
        S_c += Ql[DEN] * Ql[VEL + dim] * (S_l - Ql[VEL + dim]);
        S_c -= Qr[DEN] * Qr[VEL + dim] * (S_r - Qr[VEL + dim]);
        S_c /= (Ql[DEN] * (S_l - Ql[VEL + dim]) - Qr[DEN] * (S_r - Qr[VEL + dim]));

        const double* Q = (S_c >= 0) ? Ql : Qr;
        double S = (S_c >= 0) ? S_l : S_r;
        double U[7], Fk[7], Uc[7];
        orc_UfromQ(U, Q, gamma);
        orc_FfromQ(Fk, Q, gamma, dim);
        double A = Q[DEN] * (S - Q[VEL + dim]) / (S - S_c);
        Uc[DEN] = A;
        for (int id = 0; id < 3; ++id) Uc[VEL + id] = A * Q[VEL + id];
        Uc[VEL + dim] = A * S_c;
        Uc[PRE] = A * ((U[PRE] / Q[DEN]) + (S_c - Q[VEL + dim]) * (S_c + Q[PRE] / (Q[DEN] * (S - Q[VEL + dim]))));
        Uc[HII] = A * Q[HII];
        Uc[ADV] = A * Q[ADV];
        for (int i = 0; i < NV; ++i) F[i] = Fk[i] + S * (Uc[i] - U[i]);
    }
    return F[HII] != F[HII];
}

static int hll(double F[7], const double* Ql, const double* Qr, double a_l2, double a_r2,
               double gamma, int dim) {                                     /* :108-140 */
    double S_l, S_r;
    wave_speeds(a_l2, a_r2, Ql, Qr, dim, &S_l, &S_r);
    if (S_r <= 0) {
        orc_FfromQ(F, Qr, gamma, dim);
    } else if (S_l >= 0) {
        orc_FfromQ(F, Ql, gamma, dim);
    } else {
        double U_l[7], U_r[7], F_l[7], F_r[7];
        orc_UfromQ(U_l, Ql, gamma);
        orc_UfromQ(U_r, Qr, gamma);
        orc_FfromQ(F_l, Ql, gamma, dim);
        orc_FfromQ(F_r, Qr, gamma, dim);
        for (int i = 0; i < NV; ++i)
            F[i] = (S_r * F_l[i] - S_l * F_r[i] + S_l * S_r * (U_r[i] - U_l[i])) / (S_r - S_l);
    }
    return F[HII] != F[HII];
}

/* Eigen-ordered 3-vector helpers. */
static inline void cross3(const double a[3], const double b[3], double r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}
static inline double norm3(const double a[3]) { return sqrt(dot3(a, a)); }

/* getRotationFromA2B: Riemann.cpp:161-171. R = (I + vx) + (k*vx)*vx. */
static void rot_a2b(const double A[3], const double B[3], double R[3][3]) {
    double v[3];
    cross3(A, B, v);
    double s = norm3(v);
    double c = dot3(A, B);
    double vx[3][3] = {{0, -v[2], v[1]}, {v[2], 0, -v[0]}, {-v[1], v[0], 0}};
    double k = (1.0 - c) / (s * s);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double p = (k * vx[i][0]) * vx[0][j];
            p += (k * vx[i][1]) * vx[1][j];
            p += (k * vx[i][2]) * vx[2][j];
            R[i][j] = ((i == j ? 1.0 : 0.0) + vx[i][j]) + p;
        }
}

/* rotate: Riemann.cpp:154-159, vel <- R*vel. */
static void rotate(const double R[3][3], double* v) {
    double a = v[VEL + 0], b = v[VEL + 1], c = v[VEL + 2];
    for (int i = 0; i < 3; ++i) {
        double r = R[i][0] * a;
        r += R[i][1] * b;
        r += R[i][2] * c;
        v[VEL + i] = r;
    }
}

static int rotated_hllc(double F[7], const double* Ql, const double* Qr, double a_l2, double a_r2,
                        double gamma, int dim) {                            /* :181-244 */
    int pureHLLC = 0, pureHLL = 0;
    double d[3] = {dim == 0 ? 1.0 : 0.0, dim == 1 ? 1.0 : 0.0, dim == 2 ? 1.0 : 0.0};
    double n1[3] = {Qr[VEL + 0] - Ql[VEL + 0], Qr[VEL + 1] - Ql[VEL + 1], Qr[VEL + 2] - Ql[VEL + 2]};
    double n2[3] = {0, 0, 0};
    double t[3];

    if (n1[dim] < 0)
        for (int i = 0; i < 3; ++i) n1[i] = -1.0 * n1[i];
    if (norm3(n1) < 1.0e-6) {
        pureHLLC = 1;
    } else {
        double inv = 1.0 / norm3(n1);
        for (int i = 0; i < 3; ++i) n1[i] *= inv;
        cross3(n1, d, t);
        if (norm3(t) == 0) {
            pureHLL = 1;
        } else {
            cross3(t, n1, n2);
            if (norm3(n2) == 0) {
                pureHLL = 1;
            } else {
                if (n2[dim] < 0)
                    for (int i = 0; i < 3; ++i) n2[i] = -1.0 * n2[i];
                double inv2 = 1.0 / norm3(n2);
                for (int i = 0; i < 3; ++i) n2[i] *= inv2;
                cross3(n2, d, t);
                if (norm3(t) == 0) pureHLLC = 1;
            }
        }
    }

    int bad = 0;
    if (!pureHLLC && !pureHLL) {
        double F1[7], F2[7], QlR[7], QrR[7], R[3][3];
        memcpy(QlR, Ql, sizeof QlR);
        memcpy(QrR, Qr, sizeof QrR);
        rot_a2b(n1, d, R);
        rotate(R, QlR);
        rotate(R, QrR);
        bad |= hll(F1, QlR, QrR, a_l2, a_r2, gamma, dim);
        rot_a2b(d, n1, R);
        rotate(R, QlR);
        rotate(R, QrR);
        rotate(R, F1);
        rot_a2b(n2, d, R);
        rotate(R, QlR);
        rotate(R, QrR);
        bad |= hllc(F2, QlR, QrR, a_l2, a_r2, gamma, dim);
        rot_a2b(d, n2, R);
        rotate(R, F2);
        double w1 = fabs(dot3(d, n1)), w2 = fabs(dot3(d, n2));
        for (int iu = 0; iu < NV; ++iu) F[iu] = w1 * F1[iu] + w2 * F2[iu];
    } else if (pureHLL) {
        bad |= hll(F, Ql, Qr, a_l2, a_r2, gamma, dim);
    } else {
        bad |= hllc(F, Ql, Qr, a_l2, a_r2, gamma, dim);
    }
    return bad;
}

int orc_riemann(int solver, double F[7], const double Ql[7], const double Qr[7],
                double a_l2, double a_r2, double gamma, int dim) {
    switch (solver) {
    case ORC_HLL:  return hll(F, Ql, Qr, a_l2, a_r2, gamma, dim);
    case ORC_HLLC: return hllc(F, Ql, Qr, a_l2, a_r2, gamma, dim);
    default:       return rotated_hllc(F, Ql, Qr, a_l2, a_r2, gamma, dim);
    }
}

void orc_riemann_batch(int solver, int n, const double* Ql, const double* Qr, const int* dims,
                       double gamma, double* F, int* status) {
    for (int i = 0; i < n; ++i) {
        const double* l = Ql + 7 * i;
        const double* r = Qr + 7 * i;
        /* sound speeds as Hydro.cpp:69-71,101-102 */
        double a_l2 = gamma * l[PRE] / l[DEN];
        double a_r2 = gamma * r[PRE] / r[DEN];
        status[i] = orc_riemann(solver, F + 7 * i, l, r, a_l2, a_r2, gamma, dims[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* Whole-grid operations: reference src/Fluid/Fluid.cpp                      */
/* ------------------------------------------------------------------------ */
static inline int64_t ncells(const orc_hydro_params* p) {
    return (int64_t)p->nx * p->ny * p->nz;
}

static void set_threads(const orc_hydro_params* p) {
#ifdef _OPENMP
    omp_set_num_threads(p->nthreads > 0 ? p->nthreads : 1);
#else
    (void)p;
#endif
}

void orc_globalQfromU(const orc_hydro_params* p, const double* U, double* Q) {  /* :162-165 */
    const int64_t n = ncells(p);
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        double u[7], q[7];
        for (int v = 0; v < NV; ++v) u[v] = U[v * n + c];
        orc_QfromU(q, u, p->gamma);
        for (int v = 0; v < NV; ++v) Q[v * n + c] = q[v];
    }
}

void orc_fixPrimitives(const orc_hydro_params* p, double* Q, const double* U) {  /* :138-150 */
    const int64_t n = ncells(p);
    const double X = p->mass_frac_h;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        Q[HII * n + c] = smax(smin(Q[HII * n + c], 1.0), 0.0);
        Q[ADV * n + c] = smax(smin(Q[ADV * n + c], 1.0), 0.0);
        Q[DEN * n + c] = smax(Q[DEN * n + c], p->dfloor);
        Q[PRE * n + c] = smax(Q[PRE * n + c], p->pfloor);
        double mu_inv = X * (Q[HII * n + c] + 1.0) + (1.0 - X) * 0.25;
        double temperature = Q[PRE * n + c] / (mu_inv * p->spec_gas_const * Q[DEN * n + c]);
        if (temperature < p->tfloor)
            Q[PRE * n + c] = mu_inv * p->spec_gas_const * U[DEN * n + c] * p->tfloor;
    }
}

/* Fluid::fixSolution, Fluid.cpp:89-136.  Returns the number of cells that
 * make the reference throw (non-finite input, zero or non-finite output). */
int orc_fixSolution(const orc_hydro_params* p, double* U) {
    const int64_t n = ncells(p);
    const double X = p->mass_frac_h, g = p->gamma;
    int bad = 0;
    set_threads(p);
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t c = 0; c < n; ++c) {
        double u[7];
        for (int v = 0; v < NV; ++v) u[v] = U[v * n + c];
        if (!isfinite(u[DEN]) || !isfinite(u[PRE])) { bad += 1; continue; }
        double hii = smax(smin(u[HII] / u[DEN], 1.0), 0.0);
        double adv = smax(smin(u[ADV] / u[DEN], 1.0), 0.0);
        double v3[3];
        double den = smax(u[DEN], p->dfloor);
        for (int dim = 0; dim < 3; ++dim) v3[dim] = u[VEL + dim] / u[DEN];
        double ke = 0.0;
        for (int dim = 0; dim < 3; ++dim) ke += v3[dim] * v3[dim];
        ke *= 0.5 * u[DEN];
        double pre = (u[PRE] - ke) * (g - 1.0);
        ke *= den / u[DEN];
        if (pre < p->pfloor) pre = p->pfloor;
        double mu_inv = X * (hii + 1.0) + (1.0 - X) * 0.25;
        double temperature = pre / (mu_inv * p->spec_gas_const * den);
        if (temperature < p->tfloor) pre = mu_inv * p->spec_gas_const * den * p->tfloor;
        u[DEN] = den;
        u[PRE] = pre / (g - 1.0) + ke;
        u[HII] = hii * den;
        u[ADV] = adv * den;
        for (int dim = 0; dim < 3; ++dim) u[VEL + dim] = den * v3[dim];
        int fail = (u[DEN] == 0 || u[PRE] == 0);
        for (int v = 0; v < NV; ++v)
            if (u[v] != u[v] || isinf(u[v])) fail = 1;
        bad += fail;
        for (int v = 0; v < NV; ++v) U[v * n + c] = u[v];
    }
    return bad;
}

/* Hydrodynamics::calculateTimeStep, Hydro.cpp:77-89 (CFL literal 0.5). */
double orc_hydro_timestep(const orc_hydro_params* p, const double* Q) {
    const int64_t n = ncells(p);
    double tmin = p->dt_max;
    set_threads(p);
#pragma omp parallel
    {
        double local = p->dt_max;
#pragma omp for schedule(static)
        for (int64_t c = 0; c < n; ++c) {
            double inv_t = 0;
            double ss = sqrt(p->gamma * Q[PRE * n + c] / Q[DEN * n + c]);
            for (int dim = 0; dim < 3; ++dim)
                inv_t += (fabs(Q[(VEL + dim) * n + c]) + ss) / p->dx;
            if (inv_t == 0) inv_t = 1.0 / p->dt_max;
            local = smin(local, 0.5 / inv_t);
        }
#pragma omp critical
        tmin = smin(tmin, local);
    }
    return smin(tmin, p->dt_max);
}

/* Ghost fill for one line along a dimension: Grid::applyBCs, Grid.cpp:649-802.
 * line[g + i][v] holds cell i for i in [-ng, n+ng). */
static void fill_ghosts(double (*line)[7], int n, int ng, int dim, int bcl, int bcr) {
    const int o = ng; /* offset of cell 0 */
    for (int L = 1; L <= ng; ++L) {
        double* gh = line[o - L];
        switch (bcl) {
        case ORC_BC_PERIODIC:  memcpy(gh, line[o + n - L], sizeof(double) * 7); break;
        case ORC_BC_REFLECTING:
            memcpy(gh, line[o + L], sizeof(double) * 7);   /* Grid.cpp:765-779 */
            gh[VEL + dim] = -line[o + L][VEL + dim];
            break;
        default:
            memcpy(gh, line[o], sizeof(double) * 7);
            if (bcl == ORC_BC_OUTFLOW && line[o][VEL + dim] > 0) gh[VEL + dim] = -1.0 * line[o][VEL + dim];
            if (bcl == ORC_BC_INFLOW && line[o][VEL + dim] < 0) gh[VEL + dim] = -1.0 * line[o][VEL + dim];
        }
    }
    for (int L = 1; L <= ng; ++L) {
        double* gh = line[o + n - 1 + L];
        switch (bcr) {
        case ORC_BC_PERIODIC:  memcpy(gh, line[o + L - 1], sizeof(double) * 7); break;
        case ORC_BC_REFLECTING:
            memcpy(gh, line[o + n - 1 - L], sizeof(double) * 7);
            gh[VEL + dim] = -line[o + n - 1 - L][VEL + dim];
            break;
        default:
            memcpy(gh, line[o + n - 1], sizeof(double) * 7);
            if (bcr == ORC_BC_OUTFLOW && line[o + n - 1][VEL + dim] < 0) gh[VEL + dim] = -1.0 * line[o + n - 1][VEL + dim];
            if (bcr == ORC_BC_INFLOW && line[o + n - 1][VEL + dim] > 0) gh[VEL + dim] = -1.0 * line[o + n - 1][VEL + dim];
        }
    }
}

/* Hydrodynamics::integrate + updateSourceTerms (Hydro.cpp:31-34,44-67,91-156)
 * for one state Q.  UDOT is overwritten (the reference's UDOT is zero on
 * entry: advSolution zeroes it, Fluid.cpp:80-87). */
int orc_hydro_rhs(const orc_hydro_params* p, const double* Q, double* UDOT) {
    const int64_t n = ncells(p);
    const int N[3] = {p->nx, p->ny, p->nz};
    const int64_t stride[3] = {1, p->nx, (int64_t)p->nx * p->ny};
    const int ng = p->spatial_order + 1;
    const double g = p->gamma;
    /* Cartesian geometry, Grid.cpp:846-907: area = 1*dx*dx, vol = 1*dx*dx*dx */
    const double area = 1.0 * p->dx * p->dx;
    const double vol = 1.0 * p->dx * p->dx * p->dx;
    int bad = 0;

    memset(UDOT, 0, sizeof(double) * 7 * n);
    set_threads(p);
    for (int dim = 0; dim < 3; ++dim) {
        const int nd = N[dim];
        const int o1 = (dim + 1) % 3, o2 = (dim + 2) % 3;
        const int64_t nlines = (int64_t)N[o1] * N[o2];
#pragma omp parallel reduction(| : bad)
        {
            double (*line)[7] = malloc(sizeof(double) * 7 * (nd + 2 * ng));
            double (*QL)[7] = malloc(sizeof(double) * 7 * (nd + 2));
            double (*QR)[7] = malloc(sizeof(double) * 7 * (nd + 2));
            double (*F)[7] = malloc(sizeof(double) * 7 * (nd + 1));
#pragma omp for schedule(static)
            for (int64_t li = 0; li < nlines; ++li) {
                int a = (int)(li % N[o1]), b = (int)(li / N[o1]);
                int64_t base = a * stride[o1] + b * stride[o2];
                for (int i = 0; i < nd; ++i)
                    for (int v = 0; v < NV; ++v) line[ng + i][v] = Q[v * n + base + i * stride[dim]];
                fill_ghosts(line, nd, ng, dim, p->bc_left[dim], p->bc_right[dim]);
                /* reconstruct cells -1..nd (ghost layer 1 included), Hydro.cpp:52-67 */
                for (int i = -1; i <= nd; ++i) {
                    const double* qc = line[ng + i];
                    for (int v = 0; v < NV; ++v) {
                        if (p->spatial_order == 1) {
                            const double* ql = line[ng + i - 1];
                            const double* qr = line[ng + i + 1];
                            double dQdr = orc_limiter(p->limiter, qc[v] - ql[v], qr[v] - qc[v]);
                            QL[i + 1][v] = qc[v] - 0.5 * dQdr;
                            QR[i + 1][v] = qc[v] + 0.5 * dQdr;
                        } else {
                            QL[i + 1][v] = qc[v];
                            QR[i + 1][v] = qc[v];
                        }
                    }
                }
                /* face f between cell f-1 and cell f, Hydro.cpp:96-113 */
                for (int f = 0; f <= nd; ++f) {
                    const double* l = QR[f];      /* cell f-1 */
                    const double* r = QL[f + 1];  /* cell f   */
                    double a_l2 = g * l[PRE] / l[DEN];
                    double a_r2 = g * r[PRE] / r[DEN];
                    bad |= orc_riemann(p->solver, F[f], l, r, a_l2, a_r2, g, dim);
                }
                /* flux divergence, Hydro.cpp:125-131 */
                for (int i = 0; i < nd; ++i) {
                    int64_t c = base + i * stride[dim];
                    for (int v = 0; v < NV; ++v) {
                        double u = UDOT[v * n + c];
                        u += (area / vol) * F[i][v];
                        u -= (area / vol) * F[i + 1][v];
                        UDOT[v * n + c] = u;
                    }
                }
            }
            free(line); free(QL); free(QR); free(F);
        }
    }
    /* gravity source with GRAV = 0, Hydro.cpp:134-137 */
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        for (int i = 0; i < 3; ++i) {
            const double grav = 0.0;
            UDOT[(VEL + i) * n + c] += grav;
            UDOT[PRE * n + c] += Q[(VEL + i) * n + c] * grav;
        }
    }
    return bad;
}

static void adv_solution(const orc_hydro_params* p, double* U, double* UDOT, double dt) {
    const int64_t n = ncells(p);                                            /* Fluid.cpp:80-87 */
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c)
        for (int v = 0; v < NV; ++v) {
            U[v * n + c] += dt * UDOT[v * n + c];
            UDOT[v * n + c] = 0;
        }
}

/* Torch::hydroStep(dt, true), Torch.cpp:366-387. */
int orc_hydro_step(const orc_hydro_params* p, double* U, double* Q, double* W, double* UDOT, double dt) {
    const int64_t n = ncells(p);
    int err = 0;
    memcpy(W, U, sizeof(double) * 7 * n);                      /* globalWfromU */
    if (orc_hydro_rhs(p, Q, UDOT)) err = 1;                      /* integrate + updateSourceTerms */
    adv_solution(p, U, UDOT, dt / 2.0);
    if (orc_fixSolution(p, U) && !err) err = 2;
    orc_globalQfromU(p, U, Q);                                   /* corrector */
    memcpy(U, W, sizeof(double) * 7 * n);                        /* globalUfromW */
    if (orc_hydro_rhs(p, Q, UDOT) && !err) err = 1;
    adv_solution(p, U, UDOT, dt);
    if (orc_fixSolution(p, U) && !err) err = 2;
    return err;
}

/* Checkpointer::update, src/IO/Checkpointer.cpp:14-26. */
static double checkpoint_dt(double t, double tmax, int ncheck, int* count) {
    double delta = 100.0 / ncheck;
    double pct = smin(100.0, 100.0 * t / tmax);
    while (pct - *count * delta > delta - 1.0e-8) *count += 1;
    return ((*count + 1) * delta - pct) * tmax / 100.0;
}

/* Torch::run / fullStep / calculateTimeStep for a hydro-only run,
 * Torch.cpp:240-302, 304-339, 389-404, with the snapshots Torch::run prints
 * (DataPrinter::print2D of Fluid's Q, Torch.cpp:254, 270-281, 295-297) when
 * qprint is given: the initial fixPrimitives(QfromU(U)); at the start of
 * every step for which Checkpointer::update reports a new checkpoint, the Q
 * the previous step left (QfromU(U*) of its corrector); after the loop, if
 * the last checkpoint print was not the final one and tmax was reached.
 * Each print stores Q (7n), the checkpoint count and the time. */
static int run_hydro_impl(const orc_hydro_params* p, double* U, int nsteps, double t0, double tmax,
                          int ncheckpoints, double* dts, double* time_out, double* qprint, int max_prints,
                          int* pcount, double* ptime, int* nprint, int debug, int* steps_out, int* quit_out) {
    const int64_t n = ncells(p);
    double* Q = malloc(sizeof(double) * 7 * n);
    double* W = malloc(sizeof(double) * 7 * n);
    double* UDOT = calloc(7 * n, sizeof(double));
    double t = t0;
    int count = 0, err = 0, np = 0, final_print = 0, quitting = 0, steps = 0;
    (void)checkpoint_dt(t, tmax, ncheckpoints, &count);
#define PRINT_Q(src, cnt)                                                       \
    do {                                                                      \
        if (qprint && np < max_prints) {                                      \
            memcpy(qprint + (size_t)np * 7 * n, (src), sizeof(double) * 7 * n); \
            pcount[np] = (cnt);                                               \
            ptime[np] = t;                                                    \
            ++np;                                                             \
        }                                                                     \
    } while (0)
    if (qprint) {
        orc_globalQfromU(p, U, Q);
        orc_fixPrimitives(p, Q, U);
        PRINT_Q(Q, count);
    }
    for (int s = 0; s < nsteps && t < tmax && !quitting; ++s) {
        const int before = count;
        double dt_next = checkpoint_dt(t, tmax, ncheckpoints, &count);
        if (count != before) {
            PRINT_Q(Q, count);
            final_print = (count != ncheckpoints);
        }
        orc_globalQfromU(p, U, Q);
        orc_fixPrimitives(p, Q, U);
        /* Torch::checkValues("hydro before"), Torch.cpp:420-443 */
        int bad = 0;
        for (int64_t c = 0; c < n && !bad; ++c) {
            if (Q[DEN * n + c] == 0 || Q[PRE * n + c] == 0) bad = 1;
            for (int v = 0; v < NV; ++v)
                if (U[v * n + c] != U[v * n + c] || isinf(U[v * n + c])) bad = 1;
        }
        if (bad) { err = 3; break; }
        double dt;
        if (s == 0) {
            dt = p->dt_max * 1.0e-20;
        } else {
            dt = orc_hydro_timestep(p, Q);
            /* Torch::calculateTimeStep, Torch.cpp:321-332: the condition
             * tests the hydro percentage (all three of its terms) */
            if (debug && 100.0 * dt / tmax <= 1.0e-6) quitting = 1;
        }
        dt = smin(dt_next, dt);
        err = orc_hydro_step(p, U, Q, W, UDOT, dt);
        if (dts) dts[s] = dt;
        t += dt;
        ++steps;
        if (err) break;
    }
    /* the print after Torch::run's loop, Torch.cpp:295-297 */
    if (final_print && (!(t < tmax) || quitting)) PRINT_Q(Q, ncheckpoints);
    if (steps_out) *steps_out = steps;
    if (quit_out) *quit_out = quitting;
#undef PRINT_Q
    if (time_out) *time_out = t;
    if (nprint) *nprint = np;
    free(Q); free(W); free(UDOT);
    return err;
}

int orc_run_hydro(const orc_hydro_params* p, double* U, int nsteps, double tmax, int ncheckpoints,
                  double* dts, double* time_out, int debug, int* steps_out, int* quit_out) {
    return run_hydro_impl(p, U, nsteps, 0.0, tmax, ncheckpoints, dts, time_out, NULL, 0, NULL, NULL, NULL, debug,
                          steps_out, quit_out);
}

int orc_run_hydro_print(const orc_hydro_params* p, double* U, int nsteps, double t0, double tmax,
                        int ncheckpoints, double* dts, double* time_out, double* qprint, int max_prints, int* pcount,
                        double* ptime, int* nprint, int debug, int* steps_out, int* quit_out) {
    return run_hydro_impl(p, U, nsteps, t0, tmax, ncheckpoints, dts, time_out, qprint, max_prints, pcount, ptime,
                          nprint, debug, steps_out, quit_out);
}

/* ------------------------------------------------------------------------ */
/* Units: reference src/Torch/Converter.cpp:17-63                            */
/* ------------------------------------------------------------------------ */
void orc_converter_set_rho_pressure_time(orc_converter* c, double rho, double pressure, double time) {
    /* defaults from the Converter constructor, Converter.cpp:7-9 */
    c->M = 1.0 * 2.0e33;
    c->L = 14.6 * 3.09e18;
    c->T = 10000.0 * 3.15569e7;
    double rho_0 = (rho > 0) ? rho : c->M / (c->L * c->L * c->L);
    double pressure_0 = (pressure > 0) ? pressure : rho_0 * c->L * c->L / (c->T * c->T);
    double time_0 = (time > 0) ? time : c->T;
    c->T = time_0;
    c->M = sqrt((pressure_0 / rho_0) * pressure_0 * pressure_0) * c->T * c->T * c->T;
    c->L = sqrt(pressure_0 / rho_0) * c->T;
}

static double convert_units(const orc_converter* c, double val, double mi, double li, double ti, int to_code) {
    double MM, LL, TT;
    MM = mi >= 0 ? c->M : 1.0 / c->M;
    MM = to_code ? 1.0 / MM : MM;
    LL = li >= 0 ? c->L : 1.0 / c->L;
    LL = to_code ? 1.0 / LL : LL;
    TT = ti >= 0 ? c->T : 1.0 / c->T;
    TT = to_code ? 1.0 / TT : TT;
    int im = (int)fabs(mi), il = (int)fabs(li), it = (int)fabs(ti);
    double result = val;
    while (im >= 1 || il >= 1 || it >= 1) {
        if (im-- >= 1) result *= MM;
        if (il-- >= 1) result *= LL;
        if (it-- >= 1) result *= TT;
    }
    double left = fabs(mi) - (int)fabs(mi);
    if (left != 0) result *= pow(MM, left);
    left = fabs(li) - (int)fabs(li);
    if (left != 0) result *= pow(LL, left);
    left = fabs(ti) - (int)fabs(ti);
    if (left != 0) result *= pow(TT, left);
    return result;
}
double orc_to_code_units(const orc_converter* c, double v, double mi, double li, double ti) {
    return convert_units(c, v, mi, li, ti, 1);
}
double orc_from_code_units(const orc_converter* c, double v, double mi, double li, double ti) {
    return convert_units(c, v, mi, li, ti, 0);
}

/* ------------------------------------------------------------------------ */
/* Thermodynamics: reference src/Integrators/Thermodynamics.cpp              */
/* ------------------------------------------------------------------------ */

/* Natural cubic spline second derivatives, SplineData.cpp:15-55. */
static void spline_y2(const double* x, const double* f, int n, double* y2) {
    double u[64];
    y2[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; i++) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (f[i + 1] - f[i]) / (x[i + 1] - x[i]) - (f[i] - f[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    double qn = 0.0, un = 0.0;
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

/* SplineData::splint, SplineData.cpp:57-81. */
static double splint(const double* x, const double* f, const double* f2, int n, double x2) {
    int klo = 0, khi = n - 1;
    while (khi - klo > 1) {
        int k = (khi + klo) >> 1;
        if (x[k] > x2) khi = k;
        else klo = k;
    }
    double h = x[khi] - x[klo];
    double a = (x[khi] - x2) / h;
    double b = (x2 - x[klo]) / h;
    return a * f[klo] + b * f[khi] + ((a * a * a - a) * f2[klo] + (b * b * b - b) * f2[khi]) * (h * h) / 6.0;
}

/* LogSplineData::interpolate, SplineData.cpp:142-151. */
double orc_log_spline(const orc_thermo_params* t, double x) {
    if (x > t->spl_x[25]) return t->spl_y[25] + t->spl_max_slope * (x - t->spl_x[25]);
    if (x < t->spl_x[0]) return t->spl_y[0] + t->spl_min_slope * (x - t->spl_x[0]);
    return splint(t->spl_x, t->spl_y, t->spl_y2, 26, x);
}

/* Thermodynamics::initialise + initCollisionalExcitationHI, :24-80;
 * Constants::initialise, Constants.cpp:27-40. */
void orc_thermo_init(orc_thermo_params* t, const orc_converter* c, double thermo_hii_switch,
                     double heating_amplification, double mass_frac_h, double gamma, double pfloor,
                     int subcycling) {
    memset(t, 0, sizeof *t);
    t->hydrogen_mass = orc_to_code_units(c, 1.674e-24, 1, 0, 0);
    t->spec_gas_const = orc_to_code_units(c, 8.314462e7, 0, 2, -2);
    t->boltzmann = orc_to_code_units(c, 1.3806488e-16, 1, 2, -2);
    t->dust_xsec = orc_to_code_units(c, 5.0e-22, 0, 2, 0);
    t->pi = 3.14159265359;
    t->z0 = 5.0e-4;
    t->T1 = 33610; t->T2 = 2180; t->T3 = 28390; t->T4 = 1780;
    t->imlc = orc_to_code_units(c, 2.905e-19, 1.0, 5.0, -3.0);
    t->nmlc = orc_to_code_units(c, 4.477e-20, 1.0, 5.0, -3.0);
    t->ciec_minT = 5.0e4;
    t->ciec = orc_to_code_units(c, 3.485e-15, 1.0, 5.0, -3.0);
    t->cxhi_damp = 5.0e5;
    t->n0 = orc_to_code_units(c, 1.0e6, 0.0, -3.0, 0.0);
    t->nmc = orc_to_code_units(c, 3.981e-27, 1.0, 3.8, -3.0);
    t->fuvh_a = orc_to_code_units(c, 1.9e-26, 1.0, 4.0, -2.0);
    t->fuvh_b = orc_to_code_units(c, 1.00000, 0.0, 0.0, -1.0);
    t->fuvh_c = orc_to_code_units(c, 6.40000, 0.0, -1.0, 0.0);
    t->irh_a = orc_to_code_units(c, 7.7e-32, 1.0, 4.0, -2.0);
    t->irh_b = orc_to_code_units(c, 3.0e4, 0.0, -3.0, 0.0);
    t->crh = orc_to_code_units(c, 5.0e-27, 1.0, 2.0, -3.0);
    t->T_min_default = 100;

    static const double T[26] = {3162.2776602, 3981.0717055, 5011.8723363, 6309.5734448, 7943.2823472,
        10000.0000000, 12589.2541179, 15848.9319246, 19952.6231497, 25118.8643151,
        31622.7766017, 39810.7170553, 50118.7233627, 63095.7344480, 79432.8234724,
        100000.0000000, 125892.5411794, 158489.3192461, 199526.2314969, 251188.6431510,
        316227.7660168, 398107.1705535, 501187.2336273, 630957.3444802, 794328.2347243,
        1000000.0000000};
    static const double R[26] = {1.150800e-34, 2.312065e-31, 9.571941e-29, 1.132400e-26, 4.954502e-25,
        9.794900e-24, 1.035142e-22, 6.652732e-22, 2.870781e-21, 9.036495e-21, 2.218196e-20,
        4.456562e-20, 7.655966e-20, 1.158777e-19, 1.588547e-19, 2.013724e-19, 2.393316e-19,
        2.710192e-19, 2.944422e-19, 3.104560e-19, 3.191538e-19, 3.213661e-19, 3.191538e-19,
        3.126079e-19, 3.033891e-19, 2.917427e-19};
    for (int i = 0; i < 26; ++i) {
        t->spl_x[i] = log10(T[i]);
        t->spl_y[i] = log10(orc_to_code_units(c, R[i], 1, 5, -3));
    }
    spline_y2(t->spl_x, t->spl_y, 26, t->spl_y2);
    t->spl_min_slope = (t->spl_y[1] - t->spl_y[0]) / (t->spl_x[1] - t->spl_x[0]);
    t->spl_max_slope = (t->spl_y[25] - t->spl_y[24]) / (t->spl_x[25] - t->spl_x[24]);

    t->thermo_hii_switch = thermo_hii_switch;
    t->heating_amplification = heating_amplification;
    t->mass_frac_h = mass_frac_h;
    t->gamma = gamma;
    t->pfloor = pfloor;
    t->subcycling = subcycling;
}

/* Rate functions, Thermodynamics.cpp:113-203. */
static inline double imlc(const orc_thermo_params* t, double ne, double T) {
    return t->imlc * t->z0 * ne * ne * exp(-t->T1 / T - (t->T2 / T) * (t->T2 / T));
}
static inline double nmlc(const orc_thermo_params* t, double ne, double nn, double T) {
    return t->nmlc * t->z0 * ne * nn * exp(-t->T3 / T - (t->T4 / T) * (t->T4 / T));
}
static inline double ciec(const orc_thermo_params* t, double ne, double T) {
    if (T > t->ciec_minT) {
        double cie_rate = t->ciec * ne * ne * t->z0 * exp(-0.63 * log(T)) * (1.0 - exp(-pow(1.0e-5 * T, 1.63)));
        double smoothing = smin(1.0, (T - 5.0e4) / (2.0e4));
        return cie_rate * smoothing;
    }
    return 0;
}
static inline double nmc(const orc_thermo_params* t, double nH, double x, double T) {
    double T0 = 70.0 + 220.0 * pow(nH / t->n0, 0.2);
    return t->nmc * (1.0 - x) * (1.0 - x) * pow(nH, 1.6) * sqrt(T) * exp(-T0 / T);
}
static inline double cehi(const orc_thermo_params* t, double nH, double x, double T) {
    double rate = orc_log_spline(t, log10(T));
    return x * (1.0 - x) * nH * nH * exp((2.302585093 * rate) - ((T / t->cxhi_damp) * (T / t->cxhi_damp)));
}
static inline double fuvh(const orc_thermo_params* t, double nH, double Av, double F) {
    return t->fuvh_a * nH * F * exp(-1.9 * Av) / (t->fuvh_b + t->fuvh_c * F * exp(-1.9 * Av) / nH);
}
static inline double irh(const orc_thermo_params* t, double nH, double Av, double F) {
    return t->irh_a * nH * F * exp(-0.05 * Av) * exp(-2.0 * log(1.0 + t->irh_b / nH));
}
static inline double crh(const orc_thermo_params* t, double nH) { return t->crh * nH; }
static inline double soft_landing(double rate, double T, double T_min) {
    double result = rate;
    if (rate < 0.0) {
        if (T <= T_min) result = 0;
        else if (T <= T_min + 200) result = rate * (T - T_min) / 200;
    }
    return result;
}

void orc_thermo_rates(const orc_thermo_params* t, double nH, double x, double T, double Av, double F,
                      double out[9]) {
    double ne = nH * (x), nn = nH * (1.0 - x);
    out[0] = fuvh(t, nH, Av, F);
    out[1] = irh(t, nH, Av, F);
    out[2] = crh(t, nH);
    out[3] = imlc(t, ne, T);
    out[4] = nmlc(t, ne, nn, T);
    out[5] = cehi(t, nH, x, T);
    out[6] = ciec(t, ne, T);
    out[7] = nmc(t, nH, x, T);
    out[8] = orc_log_spline(t, log10(T));
}

/* Thermodynamics::fillHeatingArrays, :419-461 (the ray trace and F_FUV come
 * in as arrays).  HID order IMLC, NMLC, RHII, CEHI, CIEC, NMC, EUVH, FUVH,
 * IRH, CRH, TOT (Common.hpp:32-34). */
void orc_fill_heating(const orc_thermo_params* t, int64_t n, const double* Q, const double* col_den, const double* F,
                      const unsigned char* wind, double* heat, double* H) {
    for (int64_t c = 0; c < n; ++c) {
        double h[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!wind || !wind[c]) {
            double* tot = &heat[0 * n + c];
            double* rhii = &heat[1 * n + c];
            double* euvh = &heat[2 * n + c];
            if (Q[3 * n + c] < t->thermo_hii_switch) {
                *tot = 0;
                *rhii = 0;
                *euvh = 0;
            } else {
                const double den = Q[0 * n + c], pre = Q[1 * n + c], HIIFRAC = Q[2 * n + c];
                double nH = t->mass_frac_h * den / t->hydrogen_mass;
                double ne = HIIFRAC * nH;
                double nn = (1.0 - HIIFRAC) * nH;
                /* GridCell::temperature, GridCell.cpp:158-161 */
                double mu_inv = t->mass_frac_h * (HIIFRAC + 1.0) + (1.0 - t->mass_frac_h) * 0.25;
                double T = (pre / den) * (1.0 / mu_inv) / t->spec_gas_const;
                double F_FUV = F ? F[c] : 0.0;
                double tau = col_den ? col_den[c] : 0.0;
                double Av_FUV = 1.086 * t->dust_xsec * tau;
                h[7] = fuvh(t, nH, Av_FUV, F_FUV);
                h[8] = irh(t, nH, Av_FUV, F_FUV);
                h[9] = crh(t, nH);
                h[0] = -imlc(t, ne, T);
                h[1] = -nmlc(t, ne, nn, T);
                h[3] = -cehi(t, nH, HIIFRAC, T);
                h[4] = -ciec(t, ne, T);
                h[5] = -nmc(t, nH, HIIFRAC, T);
                *tot += *rhii + *euvh;
                h[2] = *rhii;
                h[6] = *euvh;
                h[10] = *tot;
            }
        }
        for (int i = 0; i < 11; ++i) H[i * n + c] = h[i];
    }
}

/* preTimeStepCalculations (:218-264) followed by integrate (:266-337) for one
 * non-wind cell. */
void orc_thermo_cell(const orc_thermo_params* t, double den, double pre, double hii, double adv,
                     double E, double T_min, double col_den, double F_FUV, double dt, double* heat_out,
                     double* rate_pre_out, double* rate_out, int64_t* nsub_out, double* p_sub_out) {
    const double X = t->mass_frac_h;
    *nsub_out = 0;
    *p_sub_out = pre;
    *heat_out = 0;
    if (adv < t->thermo_hii_switch) {
        *rate_pre_out = 0;
        *rate_out = 0;
        return;
    }
    double nH = X * den / t->hydrogen_mass;
    double HIIFRAC = hii;
    double ne = nH * (HIIFRAC);
    double nn = nH * (1.0 - HIIFRAC);
    double mu_inv0 = X * (hii + 1.0) + (1.0 - X) * 0.25;              /* Fluid.cpp:71-74 */
    double T = (pre / den) * (1.0 / mu_inv0) / t->spec_gas_const;
    double Av = 1.086 * t->dust_xsec * col_den;

    double rate = 0;
    rate += fuvh(t, nH, Av, F_FUV);
    rate += irh(t, nH, Av, F_FUV);
    rate += crh(t, nH);
    double heat = rate;
    rate -= imlc(t, ne, T);
    rate -= nmlc(t, ne, nn, T);
    rate -= cehi(t, nH, HIIFRAC, T);
    rate -= ciec(t, ne, T);
    rate -= nmc(t, nH, HIIFRAC, T);
    rate = soft_landing(rate, T, T_min);
    double RATE = t->heating_amplification * rate;
    *heat_out = heat;
    *rate_pre_out = RATE;
    if (!t->subcycling) {
        *rate_out = RATE;
        return;
    }

    double dti = fabs(0.10 * E / RATE);
    double mu_inv = X * (hii + 1.0) + (1.0 - X) * 0.25;
    double pre2temp = 1.0 / (mu_inv * t->spec_gas_const * den);
    double temp2pre = (mu_inv * t->spec_gas_const * den);
    double rate2dpre = smin(dt, dti) * (t->gamma - 1.0);
    double dpre2rate = 1.0 / rate2dpre;

    double pressure = pre + RATE * rate2dpre;
    double subcycleT = pressure * pre2temp;
    int64_t nsub = 0;
    if (pressure < t->pfloor || subcycleT < T_min) {
        double pf = smax(T_min * temp2pre, t->pfloor);
        subcycleT = pf * pre2temp;
        pressure = pf;
    }
    if (dt > dti) {
        double dtdti = dt / dti;
        /* (int) conversions as in :307; dtdti beyond INT_MAX is rejected by
         * callers (the reference's int conversion is undefined there). */
        int nsteps = dtdti - (int)dtdti > 0 ? (int)(dtdti + 1.0) : (int)(dtdti + 0.5);
        --nsteps;
        for (int i = 0; i < nsteps; ++i) {
            double subcycleRate = heat;
            subcycleRate -= imlc(t, ne, subcycleT);
            subcycleRate -= nmlc(t, ne, nn, subcycleT);
            subcycleRate -= cehi(t, nH, HIIFRAC, subcycleT);
            subcycleRate -= ciec(t, ne, subcycleT);
            subcycleRate -= nmc(t, nH, HIIFRAC, subcycleT);
            subcycleRate = t->heating_amplification * soft_landing(subcycleRate, subcycleT, T_min);
            pressure += subcycleRate * rate2dpre;
            subcycleT = pressure * pre2temp;
            if (pressure < t->pfloor || subcycleT < T_min) {
                double pf = smax(T_min * temp2pre, t->pfloor);
                subcycleT = pf * pre2temp;
                pressure = pf;
            }
        }
        nsub = nsteps;
    }
    *rate_out = (pressure - pre) * dpre2rate;
    *nsub_out = nsub;
    *p_sub_out = pressure;
}

void orc_thermo_batch(const orc_thermo_params* t, int64_t n, const double* in, double* out, int nthreads) {
#ifdef _OPENMP
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const double* r = in + 9 * i;
        double heat, rp, rate, ps;
        int64_t nsub;
        orc_thermo_cell(t, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], &heat, &rp, &rate, &nsub, &ps);
        double* o = out + 5 * i;
        o[0] = heat; o[1] = rp; o[2] = rate; o[3] = (double)nsub; o[4] = ps;
    }
}

/* ------------------------------------------------------------------------ */
/* Star geometry and the thermo column-density ray trace                       */
/* ------------------------------------------------------------------------ */
/* Cell centres sit at xc = index + 0.5 and the star at xs = cell + 0.5, in
 * cell units (Grid::buildCells, Grid.cpp:537-541; Star::initialise without
 * face snapping, Star.cpp:17-21).  3-D grids; the star inside the grid. */

/* Grid::getRayPlane, Grid.cpp:909-920: the axis of the largest |xc - xs|,
 * the first one on ties. */
static int ray_plane(const double xc[3], const double xs[3]) {
    int plane = 0;
    double m = 0;
    for (int i = 0; i < 3; ++i) {
        double a = fabs(xc[i] - xs[i]);
        if (a > m) {
            m = a;
            plane = i;
        }
    }
    return plane;
}

/* Grid::traverse3D, Grid.cpp:167-178: the cell dc[a] steps along axis
 * ax[a], or -1 off the grid. */
static int64_t step_cell(const int n[3], const int ijk[3], const int ax[3], const int dc[3]) {
    int c[3] = {ijk[0], ijk[1], ijk[2]};
    for (int a = 0; a < 3; ++a) c[ax[a]] += dc[a];
    for (int a = 0; a < 3; ++a)
        if (c[a] < 0 || c[a] >= n[a]) return -1;
    return c[0] + (int64_t)n[0] * (c[1] + (int64_t)n[1] * c[2]);
}

/* Grid::calculateNearestNeighbours (Grid.cpp:922-951), Radiation::
 * cellPathLength (Radiation.cpp:102-117) and shellVolume (:119-138) of the
 * cell (i, j, k): the four upstream neighbours and their weights, the path
 * length ds and the shell volume (code units, dx the cell width). */
void orc_star_cell_geometry(const int n[3], const int star[3], double dx, double pi, int i, int j, int k,
                            int64_t nb[4], double w[4], double* ds, double* shell_vol) {
    const int ijk[3] = {i, j, k};
    double xc[3], xs[3];
    for (int a = 0; a < 3; ++a) {
        xc[a] = (double)ijk[a] + 0.5;
        xs[a] = (double)star[a] + 0.5;
    }
    const int plane = ray_plane(xc, xs);
    const int irot[3] = {(plane + 1) % 3, (plane + 2) % 3, plane % 3};
    double d[3];
    int s[3], LR[3];
    for (int a = 0; a < 3; ++a) {
        d[a] = xc[irot[a]] - xs[irot[a]];
        s[a] = d[a] < -1.0 / 10.0 ? -1 : 1;
        LR[a] = fabs(d[a]) < 1.0 / 10.0 ? 0 : s[a];
    }
    const int o0[3] = {0, 0, -LR[2]}, o1[3] = {0, -LR[1], -LR[2]}, o2[3] = {-LR[0], 0, -LR[2]},
              o3[3] = {-LR[0], -LR[1], -LR[2]};
    nb[0] = step_cell(n, ijk, irot, o0);
    nb[1] = step_cell(n, ijk, irot, o1);
    nb[2] = step_cell(n, ijk, irot, o2);
    nb[3] = step_cell(n, ijk, irot, o3);
    const double i0 = (double)(int)xc[irot[0]], i1 = (double)(int)xc[irot[1]];
    const double ic0 = i0 - 0.5 * (s[2] * d[0] / d[2]);
    const double ic1 = i1 - 0.5 * (s[2] * d[1] / d[2]);
    const double del0 = fabs(2.0 * ic0 - 2.0 * i0 + s[0]);
    const double del1 = fabs(2.0 * ic1 - 2.0 * i1 + s[1]);
    const int f0 = fabs(d[0]) > 0.9, f1 = fabs(d[1]) > 0.9, f2 = fabs(d[2]) > 0.9;
    w[0] = f2 ? del0 * del1 : 0;
    w[1] = (f1 && f2) ? del0 * (1.0 - del1) : 0;
    w[2] = (f0 && f2) ? (1.0 - del0) * del1 : 0;
    w[3] = (f0 && f1 && f2) ? (1.0 - del0) * (1.0 - del1) : 0;
    /* cellPathLength: distance over the largest |d| (ties to z, then y) */
    double e[3];
    for (int a = 0; a < 3; ++a) e[a] = fabs(xc[a] - xs[a]);
    double denom;
    if (e[2] >= e[1] && e[2] >= e[0]) denom = e[2];
    else if (e[1] > e[2] && e[1] >= e[0]) denom = e[1];
    else denom = e[0];
    *ds = denom != 0 ? sqrt(e[0] * e[0] * dx * dx + e[1] * e[1] * dx * dx + e[2] * e[2] * dx * dx) / denom : 0.0;
    double r2 = 0;
    for (int a = 0; a < 3; ++a) r2 += (xc[a] - xs[a]) * (xc[a] - xs[a]) * dx * dx;
    if (r2 <= 100.0 * (*ds) * (*ds)) {
        double r = sqrt(r2), rmin = r - 0.5 * (*ds), rmax = r + 0.5 * (*ds);
        *shell_vol = 4.0 * pi * (rmax * rmax * rmax - rmin * rmin * rmin) / 3.0;
    } else {
        *shell_vol = 4.0 * pi * r2 * (*ds);
    }
}

/* The cells sorted by Chebyshev distance max_d |i_d - star_d| to the star
 * (a counting sort; within a shell by flat index).  Every upstream
 * neighbour of a cell is in the previous shell, so this order is causal. */
static int64_t* shell_order(const int n[3], const int star[3]) {
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
    int kmax = 0;
    for (int a = 0; a < 3; ++a) {
        int m = star[a] > n[a] - 1 - star[a] ? star[a] : n[a] - 1 - star[a];
        if (m > kmax) kmax = m;
    }
    int64_t* start = calloc((size_t)kmax + 2, sizeof(int64_t));
    int* shell = malloc(sizeof(int) * (size_t)nc);
    for (int64_t c = 0; c < nc; ++c) {
        const int ijk[3] = {(int)(c % n[0]), (int)((c / n[0]) % n[1]), (int)(c / ((int64_t)n[0] * n[1]))};
        int cheb = 0;
        for (int a = 0; a < 3; ++a) {
            int e = abs(ijk[a] - star[a]);
            if (e > cheb) cheb = e;
        }
        shell[c] = cheb;
        start[cheb + 1]++;
    }
    for (int k = 0; k <= kmax; ++k) start[k + 1] += start[k];
    int64_t* order = malloc(sizeof(int64_t) * (size_t)nc);
    for (int64_t c = 0; c < nc; ++c) order[start[shell[c]]++] = c;
    free(start);
    free(shell);
    return order;
}

void orc_star_geometry(const int n[3], const int star[3], double dx, double pi, int64_t* nb, double* w, double* ds,
                       double* shell_vol) {
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int64_t c = i + (int64_t)n[0] * (j + (int64_t)n[1] * k);
                orc_star_cell_geometry(n, star, dx, pi, i, j, k, nb + 4 * c, w + 4 * c, ds + c, shell_vol + c);
            }
}

/* Thermodynamics::rayTrace + updateColDen (Thermodynamics.cpp:339-417):
 * COL_DEN of every cell from its upstream neighbours' COL_DEN + DCOL_DEN
 * (weights divided by the neighbour's column, renormalised), DCOL_DEN =
 * (den / m_H) ds.  The reference visits the cells in its causal order
 * (Grid::buildCausal); every upstream neighbour of a cell lies one
 * Chebyshev shell closer to the star, so visiting the cells by increasing
 * shell gives the same values. */
void orc_thermo_ray_trace(const int n[3], const int star[3], const int64_t* nb, const double* w, const double* ds,
                          double hydrogen_mass, const double* den, double* col, double* dcol) {
    const int64_t ncell = (int64_t)n[0] * n[1] * n[2];
    int64_t* order = shell_order(n, star);
    {
        for (int64_t o = 0; o < ncell; ++o) {
            const int64_t c = order[o];
            const int i = (int)(c % n[0]), j = (int)((c / n[0]) % n[1]), k = (int)(c / ((int64_t)n[0] * n[1]));
            const int ijk[3] = {i, j, k};
            double dist2 = 0;
            for (int a = 0; a < 3; ++a) {
                double e = ((double)ijk[a] + 0.5) - ((double)star[a] + 0.5);
                dist2 += e * e;
            }
            if (dist2 > 0.95 * 0.95) {
                double cd[4] = {0, 0, 0, 0}, wr[4];
                for (int q = 0; q < 4; ++q) {
                    if (nb[4 * c + q] != -1) cd[q] = col[nb[4 * c + q]] + dcol[nb[4 * c + q]];
                    wr[q] = cd[q] == 0 ? 0 : w[4 * c + q] / cd[q];
                }
                double sum_w = wr[0] + wr[1] + wr[2] + wr[3];
                double v = 0.0;
                for (int q = 0; q < 4 && sum_w != 0; ++q) {
                    wr[q] = wr[q] / sum_w;
                    v += wr[q] * cd[q];
                }
                col[c] = v;
            } else {
                col[c] = 0;
            }
            dcol[c] = (den[c] / hydrogen_mass) * ds[c];
        }
    }
    free(order);
}

/* Fluid::initialiseGrid's CausalWind test (Fluid.cpp:41-54): a cell is a
 * wind cell when sum_d ((int)xc_d - p_d) (xc_d - p_d) <= R^2 (p the star's
 * cell, R the wind radius in cells) or it is the star's cell. */
void orc_star_wind(const int n[3], const int star[3], int wind_radius, unsigned char* wind) {
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int ijk[3] = {i, j, k};
                double d2 = 0;
                for (int a = 0; a < 3; ++a) d2 += (ijk[a] - star[a]) * (((double)ijk[a] + 0.5) - star[a]);
                const int here = i == star[0] && j == star[1] && k == star[2];
                wind[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)] = (d2 <= wind_radius * wind_radius) || here;
            }
}

/* The FUV flux of a star-on thermo preTimeStep (Thermodynamics.cpp:237-243,
 * fluxFUV :113-118): F = 0.5 S / (1.2e7 * 4 pi r^2), r^2 = sum_d
 * (xc_d - xs_d)^2 dx^2, and 0 at the star. */
void orc_star_fuv(const int n[3], const int star[3], double dx, double pi, double photon_rate, double* F) {
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int ijk[3] = {i, j, k};
                double r2 = 0;
                for (int a = 0; a < 3; ++a) {
                    double e = ((double)ijk[a] + 0.5) - ((double)star[a] + 0.5);
                    r2 += e * e * dx * dx;
                }
                const double q = 0.5 * photon_rate;
                F[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)] = r2 != 0 ? q / (1.2e7 * 4 * pi * r2) : 0;
            }
}

/* ------------------------------------------------------------------------ */
/* Radiation: reference src/Integrators/Radiation.cpp                        */
/* ------------------------------------------------------------------------ */

/* Radiation::initialise + initRecombinationHummer (Radiation.cpp:24-56,
 * 184-212), the Constants it reads (Constants.cpp:27-40) and the star's
 * photon energy and rate (Parameters.cpp:15-20).  Inputs in cgs. */
void orc_rad_init(orc_rad_params* r, const orc_converter* c, const double K[4], double sigma, double alphaB,
                  double tau0, double minX, double THI, double THII, int scheme, int coupling, int collisions_on,
                  double heating_amplification, double mass_frac_h, double photon_energy, double photon_rate) {
    memset(r, 0, sizeof *r);
    for (int i = 0; i < 4; ++i) r->K[i] = K[i];
    r->sigma = orc_to_code_units(c, sigma, 0, 2, 0);
    r->alphaB = orc_to_code_units(c, alphaB, 0, 3, -1);
    r->tau0 = tau0;
    r->minX = minX;
    r->THI = THI;
    r->THII = THII;
    r->scheme = scheme;
    r->coupling = coupling;
    r->collisions_on = collisions_on;
    r->heating_amplification = heating_amplification;
    r->mass_frac_h = mass_frac_h;
    r->hydrogen_mass = orc_to_code_units(c, 1.674e-24, 1, 0, 0);
    r->spec_gas_const = orc_to_code_units(c, 8.314462e7, 0, 2, -2);
    r->boltzmann = orc_to_code_units(c, 1.3806488e-16, 1, 2, -2);
    r->rydberg = orc_to_code_units(c, 13.6 * 1.60217646e-12, 1, 2, -2);
    r->voronov_A = orc_to_code_units(c, 2.91e-8, 0, 3, -1);
    r->voronov_X = 0.232;
    r->voronov_K = 0.39;
    r->voronov_P = 0.0;
    r->voronov_DE = r->rydberg / r->boltzmann;
    r->photon_energy = orc_to_code_units(c, photon_energy, 1, 2, -2);
    r->photon_rate = orc_to_code_units(c, photon_rate, 0, 0, -1);
    static const double alpha[31] = {9.283e-11, 8.823e-11, 8.361e-11, 7.898e-11, 7.435e-11, 6.973e-11, 6.512e-11,
        6.054e-11, 5.599e-11, 5.147e-11, 4.700e-11, 4.258e-11, 3.823e-11, 3.397e-11, 2.983e-11, 2.584e-11,
        2.204e-11, 1.847e-11, 1.520e-11, 1.226e-11, 9.696e-12, 7.514e-12, 5.710e-12, 4.257e-12, 3.117e-12,
        2.244e-12, 1.590e-12, 1.110e-12, 7.642e-13, 5.199e-13, 3.498e-13};
    static const double cool[31] = {8.287e-11, 7.821e-11, 7.356e-11, 6.892e-11, 6.430e-11, 5.971e-11, 5.515e-11,
        5.062e-11, 4.614e-11, 4.170e-11, 3.734e-11, 3.306e-11, 2.888e-11, 2.484e-11, 2.098e-11, 1.736e-11,
        1.402e-11, 1.103e-11, 8.442e-12, 6.279e-12, 4.539e-12, 3.192e-12, 2.185e-12, 1.458e-12, 9.484e-13,
        6.023e-13, 3.738e-13, 2.268e-13, 1.348e-13, 7.859e-14, 4.499e-14};
    for (int i = 0; i < 31; ++i) {
        /* Hummer (1994) case-B coefficients at log10 T = 1 + 0.2 i, over sqrt T */
        double T = exp(log(10.0) * (1.0 + 0.2 * (double)i));
        double sT = sqrt(T);
        r->rec_x[i] = r->cool_x[i] = T;
        r->rec_y[i] = orc_to_code_units(c, alpha[i] / sT, 0, 3, -1);
        r->cool_y[i] = orc_to_code_units(c, cool[i] / sT, 0, 3, -1);
    }
    spline_y2(r->rec_x, r->rec_y, 31, r->rec_y2);
    spline_y2(r->cool_x, r->cool_y, 31, r->cool_y2);
    /* LinearSplineData: logarithmic end slopes (SplineData.cpp:97-100) */
    r->rec_slope[0] = (log10(r->rec_y[1]) - log10(r->rec_y[0])) / (log10(r->rec_x[1]) - log10(r->rec_x[0]));
    r->rec_slope[1] = (log10(r->rec_y[30]) - log10(r->rec_y[29])) / (log10(r->rec_x[30]) - log10(r->rec_x[29]));
    r->cool_slope[0] = (log10(r->cool_y[1]) - log10(r->cool_y[0])) / (log10(r->cool_x[1]) - log10(r->cool_x[0]));
    r->cool_slope[1] =
        (log10(r->cool_y[30]) - log10(r->cool_y[29])) / (log10(r->cool_x[30]) - log10(r->cool_x[29]));
}

/* LinearSplineData::interpolate, SplineData.cpp:108-119. */
static double linear_spline(const double* x, const double* y, const double* y2, const double* slope, double v) {
    if (v > x[30]) return y[30] * pow(v / x[30], slope[1]);
    if (v < x[0]) return y[0] * pow(v / x[0], slope[0]);
    return splint(x, y, y2, 31, v);
}

double orc_rad_recombination(const orc_rad_params* r, double T) {  /* :214-218 */
    return linear_spline(r->rec_x, r->rec_y, r->rec_y2, r->rec_slope, T);
}

static double rad_recomb_cooling(const orc_rad_params* r, double nH, double x, double T) {  /* :225-228 */
    double rate = linear_spline(r->cool_x, r->cool_y, r->cool_y2, r->cool_slope, T);
    return x * x * nH * nH * r->boltzmann * T * rate;
}

static double rad_collisional(const orc_rad_params* r, double T) {  /* :236-241, Voronov (1997) */
    if (!r->collisions_on || T < 5.0e3) return 0.0;
    double U = r->voronov_DE / T;
    return r->voronov_A * (1. + r->voronov_P * sqrt(U)) * exp(r->voronov_K * log(U) - U) / (r->voronov_X + U);
}

static double rad_photoion(double nHI, double tau, double dtau, double shell_vol, double S) {  /* :243-251 */
    if (nHI == 0 || shell_vol == 0) return 0.0;
    double emT = exp(-tau), emdT = exp(-dtau);
    return S * emT * (1.0 - emdT) / (nHI * shell_vol);
}

static double rad_frac_rate(double A_pi, double A_ci, double A_rr, double nH, double x) {  /* :253-255 */
    return (1.0 - x) * (A_pi + x * nH * A_ci) - x * x * nH * A_rr;
}

static double rad_temperature(const orc_rad_params* r, double hii, double pre, double den) {  /* Fluid.cpp:72-75 */
    double mu_inv = r->mass_frac_h * (hii + 1.0) + (1.0 - r->mass_frac_h) * 0.25;
    return (pre / den) * (1.0 / mu_inv) / r->spec_gas_const;
}

/* Radiation::doric, Radiation.cpp:140-182. */
static void rad_doric(double dt, double* x_avg, double* x, double Api, double nHII_aB, double nHII_Aci) {
    const double eps = 1.0e-8;
    double xeq, inv_ti, old = *x;
    if (nHII_aB == 0.0) xeq = 1.0;
    else xeq = (Api + nHII_Aci) / (Api + nHII_Aci + nHII_aB);
    inv_ti = Api + nHII_Aci + nHII_aB;
    if (dt * inv_ti < 1.0e-8) {
        *x = old;
        *x = smax(smin(*x, 1.0), 0.0);
        if (1.0 - *x < eps) *x = 1.0 - eps;
        *x_avg = *x;
    } else {
        double e = exp(-dt * inv_ti);
        *x = xeq + (old - xeq) * e;
        *x = smax(smin(*x, 1.0), 0.0);
        if (1.0 - *x < eps) *x = 1.0 - eps;
        *x_avg = xeq + (old - xeq) * (1.0 - e) / (dt * inv_ti);
        *x_avg = smax(smin(*x_avg, 1.0), 0.0);
        if (1.0 - *x_avg < eps) *x_avg = 1.0 - eps;
    }
}

/* The first cell of the reference's CausalNonWind list (Grid::buildCausal
 * with nextCausalCoords / nextSnakeCoords, Grid.cpp:294-356,568-574; the
 * wind test of Fluid.cpp:41-54): octants in the order (+,+,+), (-,+,+),
 * (+,-,+), (-,-,+), (+,+,-), (-,+,-), (+,-,-), (-,-,-), each walked z-major,
 * then y, then x, away from the star (negative octants start one cell off
 * it).  Radiation::calculateTimeStep treats this cell specially on its
 * first call.  -1 if every cell is a wind cell. */
int64_t orc_first_nonwind_cell(const int n[3], const int star[3], int wind_radius) {
    unsigned char* wind = malloc((size_t)n[0] * n[1] * n[2]);
    orc_star_wind(n, star, wind_radius, wind);
    int64_t found = -1;
    for (int oct = 0; oct < 8 && found < 0; ++oct) {
        const int dir[3] = {(oct & 1) ? -1 : 1, (oct & 2) ? -1 : 1, (oct & 4) ? -1 : 1};
        int b[3];
        for (int a = 0; a < 3; ++a) b[a] = star[a] + (dir[a] < 0 ? -1 : 0);
        for (int k = b[2]; k >= 0 && k < n[2] && found < 0; k += dir[2])
            for (int j = b[1]; j >= 0 && j < n[1] && found < 0; j += dir[1])
                for (int i = b[0]; i >= 0 && i < n[0] && found < 0; i += dir[0]) {
                    int64_t c = i + (int64_t)n[0] * (j + (int64_t)n[1] * k);
                    if (!wind[c]) found = c;
                }
    }
    free(wind);
    return found;
}

/* Torch::initialise's Radiation::initField (Radiation.cpp:78-96): DTAU and
 * DTAU_A of every cell from the fixed primitives Q (R zero on entry). */
void orc_rad_init_field(const orc_rad_params* r, const int n[3], const int star[3], double dx, double pi,
                        const double* Q, double* R) {
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
    for (int64_t c = 0; c < nc; ++c) {
        int64_t nb[4];
        double w[4], ds, sv;
        orc_star_cell_geometry(n, star, dx, pi, (int)(c % n[0]), (int)((c / n[0]) % n[1]),
                               (int)(c / ((int64_t)n[0] * n[1])), nb, w, &ds, &sv);
        double nH = r->mass_frac_h * Q[DEN * nc + c] / r->hydrogen_mass;
        R[ORC_R_DTAU * nc + c] = ((1.0 - Q[HII * nc + c]) * nH) * r->sigma * ds;
        R[ORC_R_DTAU_A * nc + c] = ((1.0 - R[ORC_R_HII_A * nc + c]) * nH) * r->sigma * ds;
    }
}

/* Radiation::preTimeStepCalculations (Radiation.cpp:257-296) on the
 * CausalNonWind cells: R[HEAT].  Returns 1 on a NaN rate. */
static int rad_pre_step(const orc_rad_params* r, const int n[3], const int star[3], double dx, double pi,
                        const unsigned char* wind, const double* Q, double T_min, double* R) {
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
    for (int64_t c = 0; c < nc; ++c) {
        if (wind[c]) continue;
        int64_t nb[4];
        double w[4], ds, sv;
        orc_star_cell_geometry(n, star, dx, pi, (int)(c % n[0]), (int)((c / n[0]) % n[1]),
                               (int)(c / ((int64_t)n[0] * n[1])), nb, w, &ds, &sv);
        const double den = Q[DEN * nc + c], pre = Q[PRE * nc + c], hii = Q[HII * nc + c];
        double n_H = (r->mass_frac_h * den / r->hydrogen_mass);
        double excess = r->photon_energy - r->rydberg;
        double T = rad_temperature(r, hii, pre, den);
        double nHI = (1.0 - hii) * n_H;
        double A_pi = rad_photoion(nHI, R[ORC_R_TAU * nc + c], R[ORC_R_DTAU * nc + c], sv, r->photon_rate);
        double photoion = n_H * (1.0 - hii) * A_pi * excess;
        double recomb = rad_recomb_cooling(r, n_H, hii, T);
        double coll = hii * (1.0 - hii) * n_H * n_H * rad_collisional(r, T);
        double rate = photoion - recomb - coll;
        double soft = rate;
        if (T < T_min + 200 && rate < 0.0) soft = smin(0.0, soft * (T - T_min) / 200.0);
        soft *= r->heating_amplification;
        R[ORC_R_HEAT * nc + c] = soft;
        if (rate != rate) return 1;
    }
    return 0;
}

/* Radiation::calculateTimeStep (Radiation.cpp:298-351): min over the
 * CausalNonWind cells of the coupling and K1..K4 criteria.  *first is the
 * reference's function-static `once` (cleared by the call). */
static int rad_timestep(const orc_rad_params* r, const int n[3], const int star[3], double dx, double pi,
                        const unsigned char* wind, int64_t first_cell, int* first, const double* Q, const double* U,
                        const double* R, double dt_max, double* dt_out) {
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
    double dt = dt_max;
    /* the reference's causal order matters only for the `once` cell, which
     * is visited first */
    for (int64_t m = -1; m < nc; ++m) {
        const int64_t c = m < 0 ? first_cell : m;
        if (c < 0 || wind[c] || (m >= 0 && c == first_cell)) continue;
        int64_t nb[4];
        double w[4], ds, sv;
        orc_star_cell_geometry(n, star, dx, pi, (int)(c % n[0]), (int)((c / n[0]) % n[1]),
                               (int)(c / ((int64_t)n[0] * n[1])), nb, w, &ds, &sv);
        const double den = Q[DEN * nc + c], pre = Q[PRE * nc + c], hii = Q[HII * nc + c];
        double dtc = dt_max, dt1 = dt_max, dt2 = dt_max, dt3 = dt_max, dt4 = dt_max;
        if (r->coupling == ORC_COUPLING_NEQ && R[ORC_R_HEAT * nc + c] != 0)
            dtc = fabs(U[PRE * nc + c] / R[ORC_R_HEAT * nc + c]);
        if (r->K[0] != 0.0) {
            double T = rad_temperature(r, hii, pre, den);
            if (*first) {
                dt1 = r->K[0] * r->hydrogen_mass / (r->mass_frac_h * den * orc_rad_recombination(r, T));
                *first = 0;
            } else if (hii != 0) {
                dt1 = r->K[0] * r->hydrogen_mass / (r->mass_frac_h * den * hii * orc_rad_recombination(r, T));
            }
        }
        if (r->K[1] != 0.0) dt2 = dt_max;
        for (int q = 2; q < 4; ++q) {
            if (r->K[q] == 0.0) continue;
            double nH = r->mass_frac_h * den / r->hydrogen_mass;
            double nHI = (1.0 - hii) * nH;
            double A_pi = rad_photoion(nHI, R[ORC_R_TAU * nc + c], R[ORC_R_DTAU * nc + c], sv, r->photon_rate);
            double A_ci = rad_collisional(r, rad_temperature(r, hii, pre, den));
            double A_rr = orc_rad_recombination(r, rad_temperature(r, hii, pre, den));
            double fr = rad_frac_rate(A_pi, A_ci, A_rr, nH, hii);
            if (fr != 0.0) {
                if (q == 2) dt3 = r->K[2] * smax(0.05, 1.0 - hii) / fabs(fr);
                else dt4 = r->K[3] * 1.0 / fabs(fr);
            }
        }
        dt = smin(dt, smin(dt_max, smin(dtc, smin(dt1, smin(dt2, smin(dt3, dt4))))));
        if (dt == 0.0 || dt != dt || isinf(dt)) return 1;
    }
    *dt_out = dt;
    return 0;
}

/* Radiation::updateTauSC, Radiation.cpp:353-373 (avg selects TAU_A). */
static void rad_update_tau(const orc_rad_params* r, int64_t nc, int64_t c, const int64_t nb[4], const double w[4],
                           double dist2, int avg, double* R) {
    const int T_ = avg ? ORC_R_TAU_A : ORC_R_TAU, D_ = avg ? ORC_R_DTAU_A : ORC_R_DTAU;
    if (dist2 > 0.95) {
        double tau[4] = {0, 0, 0, 0}, wr[4];
        for (int i = 0; i < 4; ++i) {
            if (nb[i] != -1) tau[i] = R[T_ * nc + nb[i]] + R[D_ * nc + nb[i]];
            wr[i] = w[i] / smax(r->tau0, tau[i]);
        }
        double sum_w = wr[0] + wr[1] + wr[2] + wr[3];
        double v = 0.0;
        for (int i = 0; i < 4; ++i) {
            wr[i] = wr[i] / sum_w;
            v += wr[i] * tau[i];
        }
        R[T_ * nc + c] = v;
    } else {
        R[T_ * nc + c] = 0;
    }
}

/* Radiation::update_HIIfrac, Radiation.cpp:375-478, for a non-wind cell.
 * Q holds den, pre, hii, adv in place.  Returns 1 when the implicit
 * iteration fails. */
static int rad_update_hii(const orc_rad_params* r, int64_t nc, int64_t c, double ds, double sv, double dt, double* Q,
                          double* R) {
    const double den = Q[DEN * nc + c], pre = Q[PRE * nc + c];
    double n_H = r->mass_frac_h * den / r->hydrogen_mass;
    double A_pi = 0;
    double x = Q[HII * nc + c];
    double x_avg = x;
    double T = rad_temperature(r, Q[HII * nc + c], pre, den);
    double alphaB = orc_rad_recombination(r, T);
    double A_ci = rad_collisional(r, T);
    if (r->scheme == ORC_SCHEME_IMPLICIT || r->scheme == ORC_SCHEME_IMPLICIT2) {
        const double conv2 = 1.0e-3, conv_frac = 1.0e-5;
        double old;
        int niter = 0, miter = 0, converged = 0;
        double tau_avg = R[ORC_R_TAU_A * nc + c];
        if (r->scheme == ORC_SCHEME_IMPLICIT2) tau_avg = R[ORC_R_TAU * nc + c];
        while (!converged) {
            niter++;
            old = x_avg;
            x = Q[HII * nc + c];
            double dtau_avg = (1.0 - x_avg) * n_H * r->sigma * ds;
            double nHI = (1.0 - x_avg) * n_H;
            A_pi = rad_photoion(nHI, tau_avg, dtau_avg, sv, r->photon_rate);
            double nHII_aB = x_avg * n_H * alphaB;
            double nHII_Aci = x_avg * n_H * A_ci;
            rad_doric(dt, &x_avg, &x, A_pi, nHII_aB, nHII_Aci);
            if (fabs((x_avg - old) / x_avg) < conv2 || (x_avg < conv_frac) || A_pi == 0) converged = 1;
            if (niter > 10005) {
                miter++;
                niter = 0;
                x_avg = 0.5 * (x_avg + old);
            }
            if (miter > 5 || x != x) return 1;
        }
        R[ORC_R_HII_A * nc + c] = x_avg;
    } else {
        double tau = R[ORC_R_TAU * nc + c], dtau = R[ORC_R_DTAU * nc + c];
        double nHI = (1.0 - x) * (r->mass_frac_h * den / r->hydrogen_mass);
        A_pi = rad_photoion(nHI, tau, dtau, sv, r->photon_rate);
        double nHII_aB = x_avg * n_H * alphaB;
        double nHII_Aci = x_avg * n_H * A_ci;
        double dummy = x;
        rad_doric(dt, &x, &dummy, A_pi, nHII_aB, nHII_Aci);
    }
    Q[HII * nc + c] = x;
    if (x > 0.001) Q[ADV * nc + c] = 1;
    return 0;
}

/* Radiation::transferRadiation with a star (Radiation.cpp:569-651): wind
 * cells reset; the others, from the star outwards (each cell's upstream
 * neighbours one Chebyshev shell closer, as in the column-density ray
 * trace): TAU, TAU_A, the HII update, then DTAU, DTAU_A. */
static int rad_transfer(const orc_rad_params* r, const int n[3], const int star[3], double dx, double pi,
                        const unsigned char* wind, double dt, double* Q, double* R) {
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
    for (int64_t c = 0; c < nc; ++c) {
        if (!wind[c]) continue;
        R[ORC_R_TAU * nc + c] = 0;
        R[ORC_R_TAU_A * nc + c] = 0;
        R[ORC_R_DTAU * nc + c] = 0;
        R[ORC_R_DTAU_A * nc + c] = 0;
        Q[HII * nc + c] = 1;
        R[ORC_R_HII_A * nc + c] = 1;
    }
    int64_t* order = shell_order(n, star);
    int err = 0;
    for (int64_t o = 0; o < nc && !err; ++o) {
            const int64_t c = order[o];
            if (wind[c]) continue;
            const int ijk[3] = {(int)(c % n[0]), (int)((c / n[0]) % n[1]), (int)(c / ((int64_t)n[0] * n[1]))};
            double dist2 = 0;
            for (int a = 0; a < 3; ++a) {
                double d = ((double)ijk[a] + 0.5) - ((double)star[a] + 0.5);
                dist2 += d * d;
            }
            int64_t nb[4];
            double w[4], ds, sv;
            orc_star_cell_geometry(n, star, dx, pi, ijk[0], ijk[1], ijk[2], nb, w, &ds, &sv);
            rad_update_tau(r, nc, c, nb, w, dist2, 0, R);
            rad_update_tau(r, nc, c, nb, w, dist2, 1, R);
            if (rad_update_hii(r, nc, c, ds, sv, dt, Q, R)) {
                err = 1;
                break;
            }
            double nH = r->mass_frac_h * Q[DEN * nc + c] / r->hydrogen_mass;
            R[ORC_R_DTAU * nc + c] = ((1.0 - Q[HII * nc + c]) * nH) * r->sigma * ds;
            R[ORC_R_DTAU_A * nc + c] = ((1.0 - R[ORC_R_HII_A * nc + c]) * nH) * r->sigma * ds;
    }
    free(order);
    return err;
}

/* Radiation::transferRadiation2 with a star (Radiation.cpp:510-567), the
 * integrate of the implicit2 and explicit schemes: Radiation::rayTrace
 * first (:480-508: wind cells reset; the others, from the star outwards,
 * TAU from the upstream neighbours, then DTAU and DTAU_A from the HII
 * fractions before the update), then the HII update of every non-wind
 * cell (independent cells). */
static int rad_transfer2(const orc_rad_params* r, const int n[3], const int star[3], double dx, double pi,
                         const unsigned char* wind, double dt, double* Q, double* R) {
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
    for (int64_t c = 0; c < nc; ++c) {
        if (!wind[c]) continue;
        R[ORC_R_TAU * nc + c] = 0;
        R[ORC_R_TAU_A * nc + c] = 0;
        R[ORC_R_DTAU * nc + c] = 0;
        R[ORC_R_DTAU_A * nc + c] = 0;
        Q[HII * nc + c] = 1;
        R[ORC_R_HII_A * nc + c] = 1;
    }
    int64_t* order = shell_order(n, star);
    int err = 0;
    for (int64_t o = 0; o < nc && !err; ++o) {
            const int64_t c = order[o];
            if (wind[c]) continue;
            const int ijk[3] = {(int)(c % n[0]), (int)((c / n[0]) % n[1]), (int)(c / ((int64_t)n[0] * n[1]))};
            double dist2 = 0;
            for (int a = 0; a < 3; ++a) {
                double d = ((double)ijk[a] + 0.5) - ((double)star[a] + 0.5);
                dist2 += d * d;
            }
            int64_t nb[4];
            double w[4], ds, sv;
            orc_star_cell_geometry(n, star, dx, pi, ijk[0], ijk[1], ijk[2], nb, w, &ds, &sv);
            rad_update_tau(r, nc, c, nb, w, dist2, 0, R);
            double nH = r->mass_frac_h * Q[DEN * nc + c] / r->hydrogen_mass;
            R[ORC_R_DTAU * nc + c] = ((1.0 - Q[HII * nc + c]) * nH) * r->sigma * ds;
            R[ORC_R_DTAU_A * nc + c] = ((1.0 - R[ORC_R_HII_A * nc + c]) * nH) * r->sigma * ds;
    }
    free(order);
    (void)err;
    for (int64_t c = 0; c < nc; ++c) {
        if (wind[c]) continue;
        int64_t nb[4];
        double w[4], ds, sv;
        orc_star_cell_geometry(n, star, dx, pi, (int)(c % n[0]), (int)((c / n[0]) % n[1]),
                               (int)(c / ((int64_t)n[0] * n[1])), nb, w, &ds, &sv);
        if (rad_update_hii(r, nc, c, ds, sv, dt, Q, R)) return 1;
    }
    return 0;
}

/* Radiation::updateSourceTerms (Radiation.cpp:653-680) + Fluid::advSolution
 * + fixSolution for a star-on run. */
static int rad_update_source(const orc_rad_params* r, const orc_hydro_params* p, double dt, const double* Q,
                             const double* R, double* U) {
    const int64_t nc = ncells(p);
    double* UDOT = calloc(7 * (size_t)nc, sizeof(double));
    for (int64_t c = 0; c < nc; ++c) {
        const double den = Q[DEN * nc + c];
        if (r->coupling == ORC_COUPLING_TTI) {
            double x = Q[HII * nc + c];
            double mu_inv = r->mass_frac_h * (x + 1.0) + (1.0 - r->mass_frac_h) * 0.25;
            double T_new = r->THI + (r->THII - r->THI) * x;
            double pre_new = r->spec_gas_const * mu_inv * den * T_new;
            double ke = 0.0;
            for (int d = 0; d < 3; ++d) ke += 0.5 * U[(VEL + d) * nc + c] * U[(VEL + d) * nc + c] / U[DEN * nc + c];
            double E_new = pre_new / (p->gamma - 1.0) + ke;
            UDOT[PRE * nc + c] += (E_new - U[PRE * nc + c]) / dt;
        } else if (r->coupling == ORC_COUPLING_NEQ) {
            UDOT[PRE * nc + c] += R[ORC_R_HEAT * nc + c];
        }
        UDOT[HII * nc + c] += (Q[HII * nc + c] * den - U[HII * nc + c]) / dt;
        UDOT[ADV * nc + c] += (Q[ADV * nc + c] * den - U[ADV * nc + c]) / dt;
    }
    adv_solution(p, U, UDOT, dt);
    free(UDOT);
    return orc_fixSolution(p, U) ? 2 : 0;
}

static int rad_check_values(const orc_hydro_params* p, const double* U, const double* Q) {  /* Torch.cpp:420-443 */
    const int64_t nc = ncells(p);
    for (int64_t c = 0; c < nc; ++c) {
        if (Q[DEN * nc + c] == 0 || Q[PRE * nc + c] == 0) return 1;
        for (int v = 0; v < NV; ++v)
            if (U[v * nc + c] != U[v * nc + c] || isinf(U[v * nc + c])) return 1;
    }
    return 0;
}

/* The radiation-only sequence of the reference harness's radiation mode:
 * per step Q = fixPrimitives(QfromU(U)), preTimeStepCalculations, dt =
 * min(dt_max, calculateTimeStep), then Torch::subStep(dt, true) and
 * subStep(dt, false) of the radiation component (Torch.cpp:352-364).  U
 * and R (6 n, RID order) are updated; Qout (2 n) receives Q[HII], Q[ADV]
 * after the last step; dts the dt of each step.  *first as rad_timestep.
 * Returns 0, or 1 (NaN heating / timestep / implicit iteration), 2
 * (fixSolution), 3 (checkValues). */
int orc_rad_run(const orc_rad_params* r, const orc_hydro_params* p, const int star[3], int wind_radius,
                double T_min, int* first, double* U, double* R, double* Qout, int nsteps, double* dts) {
    const int n[3] = {p->nx, p->ny, p->nz};
    const int64_t nc = ncells(p);
    double* Q = malloc(sizeof(double) * 7 * nc);
    unsigned char* wind = malloc((size_t)nc);
    orc_star_wind(n, star, wind_radius, wind);
    const int64_t first_cell = orc_first_nonwind_cell(n, star, wind_radius);
    const double pi = 3.14159265359;
    int err = 0;
    for (int s = 0; s < nsteps && !err; ++s) {
        orc_globalQfromU(p, U, Q);
        orc_fixPrimitives(p, Q, U);
        if ((err = rad_pre_step(r, n, star, p->dx, pi, wind, Q, T_min, R))) break;
        double dt;
        if ((err = rad_timestep(r, n, star, p->dx, pi, wind, first_cell, first, Q, U, R, p->dt_max, &dt))) break;
        dt = smin(p->dt_max, dt);
        dts[s] = dt;
        for (int sub = 0; sub < 2 && !err; ++sub) {
            if (rad_check_values(p, U, Q)) { err = 3; break; }
            if (sub == 1) {
                orc_globalQfromU(p, U, Q);
                orc_fixPrimitives(p, Q, U);
                if ((err = rad_pre_step(r, n, star, p->dx, pi, wind, Q, T_min, R))) break;
            }
            /* Radiation::integrate (Radiation.cpp:58-63) */
            if (r->scheme == ORC_SCHEME_IMPLICIT) err = rad_transfer(r, n, star, p->dx, pi, wind, dt, Q, R);
            else err = rad_transfer2(r, n, star, p->dx, pi, wind, dt, Q, R);
            if (err) break;
            if ((err = rad_update_source(r, p, dt, Q, R, U))) break;
            /* checkValues(" after") sees U and the Q the transfer left */
            if (rad_check_values(p, U, Q)) err = 3;
        }
    }
    for (int64_t c = 0; c < nc; ++c) {
        Qout[c] = Q[HII * nc + c];
        Qout[nc + c] = Q[ADV * nc + c];
    }
    free(Q);
    free(wind);
    return err;
}
