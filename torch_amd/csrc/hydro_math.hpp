// hydro_math.hpp -- per-face and per-cell hydrodynamics arithmetic for the
// CDNA4 stage kernels, templated on the working precision R (double for the
// parity path, float for the mixed-precision flux path).
//
// For R = double every expression keeps the operand order of the reference
// (hgsteggles/TORCH) so that, compiled with -ffp-contract=off, results are
// bitwise equal to its x86-64 FP64 build:
//   slope limiters         src/Integrators/SlopeLimiter.cpp:263-311
//   UfromQ/QfromU/FfromQ   src/Fluid/GridCell.cpp:163-213
//   wave speeds            src/Integrators/Riemann.cpp:27-42
//   HLLC / HLL             src/Integrators/Riemann.cpp:52-104, 108-140
//   rotated hybrid solver  src/Integrators/Riemann.cpp:154-244 with Eigen
//                          3.2.1 fixed-size orders: cross (a1b2-a2b1, ...),
//                          dot x0+(x1+x2), mat*vec (l0r0+l1r1)+l2r2,
//                          normalize x*(1/|x|).
// The face normal DIM is a template parameter so that every component index
// is resolved at compile time (no dynamically indexed register arrays).
#pragma once

#include <hip/hip_runtime.h>

namespace tg {

enum { DEN = 0, PRE = 1, HII = 2, ADV = 3, VEL = 4, NV = 7 };
enum { SOLVER_HLL = 0, SOLVER_HLLC = 1, SOLVER_ROTATED = 2 };
enum { LIM_MC = 0, LIM_SUPERBEE = 1, LIM_MINMOD = 2, LIM_MAXMOD = 3, LIM_LEER = 4, LIM_OSPRE = 5, LIM_ALBADA = 6 };
enum { BC_FREE = 0, BC_REFLECTING = 1, BC_OUTFLOW = 2, BC_INFLOW = 3, BC_PERIODIC = 4 };

// Gas constants in working precision; g = gamma, gm1 = gamma - 1.0 and
// g2 = gamma / (gamma - 1.0) are rounded once on the host exactly as the
// reference rounds them inside each call.
template <typename R>
struct GasConsts {
    R g, gm1, g2;
};

template <typename R>
__device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }  // std::min
template <typename R>
__device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }  // std::max

template <typename R>
__device__ __forceinline__ R minmod(R a, R b) {
    if (a * b > R(0)) return fabs(a) < fabs(b) ? a : b;
    return R(0);
}
template <typename R>
__device__ __forceinline__ R maxmod(R a, R b) {
    if (a * b > R(0)) return fabs(a) > fabs(b) ? a : b;
    return R(0);
}

template <typename R>
__device__ __forceinline__ R limiter(int lim, R a, R b) {
    if (lim == LIM_ALBADA) return a * b <= R(1.0e-30) ? R(0) : a * b * (a + b) / (a * a + b * b);
    switch (lim) {
    case LIM_MC:       return minmod((a + b) / R(2.0), minmod(R(2.0) * a, R(2.0) * b));
    case LIM_SUPERBEE: return maxmod(minmod(b, R(2.0) * a), minmod(R(2.0) * b, a));
    case LIM_MINMOD:   return minmod(a, b);
    case LIM_MAXMOD:   return maxmod(a, b);
    case LIM_LEER: {
        if (b == R(0)) return R(0);
        R r = fabs(a / b);
        return (a + b * r) / (R(1.0) + r);
    }
    default: /* LIM_OSPRE */
        if (a * b <= R(1.0e-30)) return R(0);
        return R(1.5) * b * (a * a + a * b) / (a * a + a * b + b * b);
    }
}

template <typename R>
__device__ __forceinline__ void UfromQ(R u[7], const R q[7], const GasConsts<R>& c) {
    R ke = R(0);
    u[DEN] = q[DEN];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        u[VEL + d] = q[VEL + d] * q[DEN];
        ke += R(0.5) * q[DEN] * q[VEL + d] * q[VEL + d];
    }
    u[PRE] = q[PRE] / c.gm1 + ke;
    u[HII] = q[HII] * q[DEN];
    u[ADV] = q[ADV] * q[DEN];
}

// FP64 primitive recovery (kept in FP64 in both precision modes).
__device__ __forceinline__ void QfromU(double q[7], const double u[7], double gm1) {
    q[DEN] = u[DEN];
    double ke = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q[VEL + d] = u[VEL + d] / u[DEN];
        ke += 0.5 * u[VEL + d] * u[VEL + d] / u[DEN];
    }
    q[PRE] = (u[PRE] - ke) * gm1;
    q[HII] = u[HII] / u[DEN];
    q[ADV] = u[ADV] / u[DEN];
}

template <typename R, int DIM>
__device__ __forceinline__ void FfromQ(R f[7], const R q[7], const GasConsts<R>& c) {
    f[DEN] = q[DEN] * q[VEL + DIM];
    R ke = R(0);
#pragma unroll
    for (int id = 0; id < 3; ++id) {
        ke += q[VEL + id] * q[VEL + id];
        f[VEL + (DIM + id) % 3] = q[DEN] * q[VEL + (DIM + id) % 3] * q[VEL + DIM];
    }
    ke *= R(0.5) * q[DEN];
    f[VEL + DIM] += q[PRE];
    f[PRE] = q[VEL + DIM] * (c.g2 * q[PRE] + ke);
    f[HII] = q[DEN] * q[VEL + DIM] * q[HII];
    f[ADV] = q[DEN] * q[VEL + DIM] * q[ADV];
}

template <typename R, int DIM>
__device__ __forceinline__ void wave_speeds(R a_l2, R a_r2, const R* Ql, const R* Qr, R& S_l, R& S_r) {
    R sqrtrho_l = sqrt(Ql[DEN]);
    R sqrtrho_r = sqrt(Qr[DEN]);
    R u_tilde = (sqrtrho_l * Ql[VEL + DIM] + sqrtrho_r * Qr[VEL + DIM]) / (sqrtrho_l + sqrtrho_r);
    R nu2 = R(0.5) * sqrtrho_r * sqrtrho_l / ((sqrtrho_l + sqrtrho_r) * (sqrtrho_l + sqrtrho_r));
    R dsqrd = (sqrtrho_l * a_l2 + sqrtrho_r * a_r2) / (sqrtrho_l + sqrtrho_r);
    dsqrd += nu2 * (Qr[VEL + DIM] - Ql[VEL + DIM]) * (Qr[VEL + DIM] - Ql[VEL + DIM]);
    R d = sqrt(dsqrd);
    S_l = u_tilde - d;
    S_r = u_tilde + d;
}

// HartenLaxLeerContactSolver::solve.  Returns true when the HII flux is NaN
// (the reference throws there).
template <typename R, int DIM>
__device__ __forceinline__ bool hllc(R F[7], const R* Ql, const R* Qr, R a_l2, R a_r2, const GasConsts<R>& c) {
    R S_l, S_r;
    wave_speeds<R, DIM>(a_l2, a_r2, Ql, Qr, S_l, S_r);
    if (S_l >= R(0)) {
        FfromQ<R, DIM>(F, Ql, c);
    } else if (S_r <= R(0)) {
        FfromQ<R, DIM>(F, Qr, c);
    } else {
        R S_c = Qr[PRE] - Ql[PRE];
        S_c += Ql[DEN] * Ql[VEL + DIM] * (S_l - Ql[VEL + DIM]);
        S_c -= Qr[DEN] * Qr[VEL + DIM] * (S_r - Qr[VEL + DIM]);
        S_c /= (Ql[DEN] * (S_l - Ql[VEL + DIM]) - Qr[DEN] * (S_r - Qr[VEL + DIM]));
        const bool left = S_c >= R(0);
        R Q[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) Q[i] = left ? Ql[i] : Qr[i];
        const R S = left ? S_l : S_r;
        R U[7], Fk[7], Uc[7];
        UfromQ<R>(U, Q, c);
        FfromQ<R, DIM>(Fk, Q, c);
        R A = Q[DEN] * (S - Q[VEL + DIM]) / (S - S_c);
        Uc[DEN] = A;
#pragma unroll
        for (int id = 0; id < 3; ++id) Uc[VEL + id] = A * Q[VEL + id];
        Uc[VEL + DIM] = A * S_c;
        Uc[PRE] = A * ((U[PRE] / Q[DEN]) + (S_c - Q[VEL + DIM]) * (S_c + Q[PRE] / (Q[DEN] * (S - Q[VEL + DIM]))));
        Uc[HII] = A * Q[HII];
        Uc[ADV] = A * Q[ADV];
#pragma unroll
        for (int i = 0; i < 7; ++i) F[i] = Fk[i] + S * (Uc[i] - U[i]);
    }
    return F[HII] != F[HII];
}

// HartenLaxLeerSolver::solve.
template <typename R, int DIM>
__device__ __forceinline__ bool hll(R F[7], const R* Ql, const R* Qr, R a_l2, R a_r2, const GasConsts<R>& c) {
    R S_l, S_r;
    wave_speeds<R, DIM>(a_l2, a_r2, Ql, Qr, S_l, S_r);
    if (S_r <= R(0)) {
        FfromQ<R, DIM>(F, Qr, c);
    } else if (S_l >= R(0)) {
        FfromQ<R, DIM>(F, Ql, c);
    } else {
        R U_l[7], U_r[7], F_l[7], F_r[7];
        UfromQ<R>(U_l, Ql, c);
        UfromQ<R>(U_r, Qr, c);
        FfromQ<R, DIM>(F_l, Ql, c);
        FfromQ<R, DIM>(F_r, Qr, c);
#pragma unroll
        for (int i = 0; i < 7; ++i)
            F[i] = (S_r * F_l[i] - S_l * F_r[i] + S_l * S_r * (U_r[i] - U_l[i])) / (S_r - S_l);
    }
    return F[HII] != F[HII];
}

template <typename R>
__device__ __forceinline__ void cross3(const R a[3], const R b[3], R r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
template <typename R>
__device__ __forceinline__ R dot3(const R a[3], const R b[3]) {
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}
template <typename R>
__device__ __forceinline__ R norm3(const R a[3]) { return sqrt(dot3(a, a)); }

// getRotationFromA2B: R = (I + [v]x) + (k [v]x) [v]x with v = A x B,
// k = (1 - A.B) / |v|^2.
template <typename R>
__device__ __forceinline__ void rot_a2b(const R A[3], const R B[3], R M[3][3]) {
    R v[3];
    cross3(A, B, v);
    R s = norm3(v);
    R cc = dot3(A, B);
    const R vx[3][3] = {{R(0), -v[2], v[1]}, {v[2], R(0), -v[0]}, {-v[1], v[0], R(0)}};
    R k = (R(1.0) - cc) / (s * s);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            R p = (k * vx[i][0]) * vx[0][j];
            p += (k * vx[i][1]) * vx[1][j];
            p += (k * vx[i][2]) * vx[2][j];
            M[i][j] = ((i == j ? R(1.0) : R(0.0)) + vx[i][j]) + p;
        }
}

template <typename R>
__device__ __forceinline__ void rotate(const R M[3][3], R* v) {
    const R a = v[VEL + 0], b = v[VEL + 1], cz = v[VEL + 2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R r = M[i][0] * a;
        r += M[i][1] * b;
        r += M[i][2] * cz;
        v[VEL + i] = r;
    }
}

// RotatedHartenLaxLeerSolver::solve (Nishikawa & Kitamura 2008).
template <typename R, int DIM>
__device__ __forceinline__ bool rotated_hllc(R F[7], const R* Ql, const R* Qr, R a_l2, R a_r2, const GasConsts<R>& c) {
    bool pureHLLC = false, pureHLL = false;
    const R d[3] = {DIM == 0 ? R(1) : R(0), DIM == 1 ? R(1) : R(0), DIM == 2 ? R(1) : R(0)};
    R n1[3] = {Qr[VEL + 0] - Ql[VEL + 0], Qr[VEL + 1] - Ql[VEL + 1], Qr[VEL + 2] - Ql[VEL + 2]};
    R n2[3] = {R(0), R(0), R(0)};
    R t[3];
    if (n1[DIM] < R(0)) {
#pragma unroll
        for (int i = 0; i < 3; ++i) n1[i] = R(-1.0) * n1[i];
    }
    if (norm3(n1) < R(1.0e-6)) {
        pureHLLC = true;
    } else {
        const R inv = R(1.0) / norm3(n1);
#pragma unroll
        for (int i = 0; i < 3; ++i) n1[i] *= inv;
        cross3(n1, d, t);
        if (norm3(t) == R(0)) {
            pureHLL = true;
        } else {
            cross3(t, n1, n2);
            if (norm3(n2) == R(0)) {
                pureHLL = true;
            } else {
                if (n2[DIM] < R(0)) {
#pragma unroll
                    for (int i = 0; i < 3; ++i) n2[i] = R(-1.0) * n2[i];
                }
                const R inv2 = R(1.0) / norm3(n2);
#pragma unroll
                for (int i = 0; i < 3; ++i) n2[i] *= inv2;
                cross3(n2, d, t);
                if (norm3(t) == R(0)) pureHLLC = true;
            }
        }
    }
    bool bad = false;
    if (!pureHLLC && !pureHLL) {
        R F1[7], F2[7], QlR[7], QrR[7], M[3][3];
#pragma unroll
        for (int i = 0; i < 7; ++i) { QlR[i] = Ql[i]; QrR[i] = Qr[i]; }
        rot_a2b(n1, d, M);
        rotate(M, QlR);
        rotate(M, QrR);
        bad |= hll<R, DIM>(F1, QlR, QrR, a_l2, a_r2, c);
        rot_a2b(d, n1, M);
        rotate(M, QlR);
        rotate(M, QrR);
        rotate(M, F1);
        rot_a2b(n2, d, M);
        rotate(M, QlR);
        rotate(M, QrR);
        bad |= hllc<R, DIM>(F2, QlR, QrR, a_l2, a_r2, c);
        rot_a2b(d, n2, M);
        rotate(M, F2);
        const R w1 = fabs(dot3(d, n1)), w2 = fabs(dot3(d, n2));
#pragma unroll
        for (int i = 0; i < 7; ++i) F[i] = w1 * F1[i] + w2 * F2[i];
    } else if (pureHLL) {
        bad |= hll<R, DIM>(F, Ql, Qr, a_l2, a_r2, c);
    } else {
        bad |= hllc<R, DIM>(F, Ql, Qr, a_l2, a_r2, c);
    }
    return bad;
}

template <int SOLVER, typename R, int DIM>
__device__ __forceinline__ bool riemann(R F[7], const R* Ql, const R* Qr, R a_l2, R a_r2, const GasConsts<R>& c) {
    if constexpr (SOLVER == SOLVER_HLL) return hll<R, DIM>(F, Ql, Qr, a_l2, a_r2, c);
    else if constexpr (SOLVER == SOLVER_HLLC) return hllc<R, DIM>(F, Ql, Qr, a_l2, a_r2, c);
    else return rotated_hllc<R, DIM>(F, Ql, Qr, a_l2, a_r2, c);
}

}  // namespace tg
