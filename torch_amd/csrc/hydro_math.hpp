// hydro_math.hpp -- per-face and per-cell hydrodynamics arithmetic for the
// CDNA4 stage kernels, templated on the working precision R (double for the
// parity path, float for the mixed-precision flux path).
//
// For R = double every expression keeps the operand order of the reference
// (hgsteggles/TORCH) so that, compiled with -ffp-contract=off, results are
// bitwise equal to its x86-64 FP64 build:
//   slope limiters         src/Integrators/SlopeLimiter.cpp:7-54
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
// (the reference throws there).  Written branch-free: the upwind side is
// selected first and F(Q_side) is the reference's supersonic flux and the
// Fk of its subsonic star-state flux alike, so every lane runs one FfromQ;
// the values are those of the reference's branches.
template <typename R, int DIM>
__device__ __forceinline__ bool hllc(R F[7], const R* Ql, const R* Qr, R a_l2, R a_r2, const GasConsts<R>& c) {
    R S_l, S_r;
    wave_speeds<R, DIM>(a_l2, a_r2, Ql, Qr, S_l, S_r);
    R S_c = Qr[PRE] - Ql[PRE];
    S_c += Ql[DEN] * Ql[VEL + DIM] * (S_l - Ql[VEL + DIM]);
    S_c -= Qr[DEN] * Qr[VEL + DIM] * (S_r - Qr[VEL + DIM]);
    S_c /= (Ql[DEN] * (S_l - Ql[VEL + DIM]) - Qr[DEN] * (S_r - Qr[VEL + DIM]));
    const bool sup_l = S_l >= R(0);
    const bool sup_r = !sup_l && S_r <= R(0);
    const bool left = sup_l || (!sup_r && S_c >= R(0));
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
    const bool sub = !sup_l && !sup_r;
#pragma unroll
    for (int i = 0; i < 7; ++i) F[i] = sub ? Fk[i] + S * (Uc[i] - U[i]) : Fk[i];
    return F[HII] != F[HII];
}

// HartenLaxLeerSolver::solve, branch-free (both one-sided fluxes are needed
// by the subsonic case; the three candidates are selected per lane).
template <typename R, int DIM>
__device__ __forceinline__ bool hll(R F[7], const R* Ql, const R* Qr, R a_l2, R a_r2, const GasConsts<R>& c) {
    R S_l, S_r;
    wave_speeds<R, DIM>(a_l2, a_r2, Ql, Qr, S_l, S_r);
    R U_l[7], U_r[7], F_l[7], F_r[7];
    UfromQ<R>(U_l, Ql, c);
    UfromQ<R>(U_r, Qr, c);
    FfromQ<R, DIM>(F_l, Ql, c);
    FfromQ<R, DIM>(F_r, Qr, c);
    const bool right = S_r <= R(0);
    const bool left = !right && S_l >= R(0);
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const R hl = (S_r * F_l[i] - S_l * F_r[i] + S_l * S_r * (U_r[i] - U_l[i])) / (S_r - S_l);
        F[i] = right ? F_r[i] : (left ? F_l[i] : hl);
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
    // One HLL and one HLLC solve per face for every lane: on the hybrid path
    // they take the rotated states, on the pure paths the unrotated ones (the
    // reference's single solve; the other result is discarded), so a wave
    // with mixed paths runs the solvers once instead of once per path.
    const bool hybrid = !pureHLLC && !pureHLL;
    R F1[7], F2[7], QlR[7], QrR[7], Qa[7], Qb[7], M[3][3];
#pragma unroll
    for (int i = 0; i < 7; ++i) { QlR[i] = Ql[i]; QrR[i] = Qr[i]; }
    rot_a2b(n1, d, M);
    rotate(M, QlR);
    rotate(M, QrR);
#pragma unroll
    for (int i = 0; i < 7; ++i) { Qa[i] = hybrid ? QlR[i] : Ql[i]; Qb[i] = hybrid ? QrR[i] : Qr[i]; }
    const bool bad1 = hll<R, DIM>(F1, Qa, Qb, a_l2, a_r2, c);
    rot_a2b(d, n1, M);
    rotate(M, QlR);
    rotate(M, QrR);
    // A = w1 * (M F1) on the hybrid path (the first term of the reference's
    // w1 F1 + w2 F2), F1 itself on the pure-HLL path
    const R w1 = fabs(dot3(d, n1));
    {
        R Fr[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) Fr[i] = F1[i];
        rotate(M, Fr);
#pragma unroll
        for (int i = 0; i < 7; ++i) F1[i] = hybrid ? w1 * Fr[i] : F1[i];
    }
    rot_a2b(n2, d, M);
    rotate(M, QlR);
    rotate(M, QrR);
#pragma unroll
    for (int i = 0; i < 7; ++i) { Qa[i] = hybrid ? QlR[i] : Ql[i]; Qb[i] = hybrid ? QrR[i] : Qr[i]; }
    const bool bad2 = hllc<R, DIM>(F2, Qa, Qb, a_l2, a_r2, c);
    rot_a2b(d, n2, M);
    const R w2 = fabs(dot3(d, n2));
    R Fr[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) Fr[i] = F2[i];
    rotate(M, Fr);
#pragma unroll
    for (int i = 0; i < 7; ++i) F[i] = hybrid ? F1[i] + w2 * Fr[i] : (pureHLL ? F1[i] : F2[i]);
    return hybrid ? (bad1 || bad2) : (pureHLL ? bad1 : bad2);
}

template <int SOLVER, typename R, int DIM>
__device__ __forceinline__ bool riemann(R F[7], const R* Ql, const R* Qr, R a_l2, R a_r2, const GasConsts<R>& c) {
    if constexpr (SOLVER == SOLVER_HLL) return hll<R, DIM>(F, Ql, Qr, a_l2, a_r2, c);
    else if constexpr (SOLVER == SOLVER_HLLC) return hllc<R, DIM>(F, Ql, Qr, a_l2, a_r2, c);
    else return rotated_hllc<R, DIM>(F, Ql, Qr, a_l2, a_r2, c);
}

// ---------------------------------------------------------------------------
// Mixed-precision face flux (float).
//
// The same solvers, reorganised for throughput rather than bitwise parity:
//  * reciprocals and square roots are the hardware's (v_rcp_f32, v_sqrt_f32,
//    ~1 ulp), each shared wherever the reference divides twice by the same
//    quantity;
//  * the rotated hybrid solver is evaluated without rotation matrices.
//    Rotating both states so that a unit normal n maps onto the face axis,
//    solving there and rotating the flux back (Riemann.cpp:154-244) is the
//    same as solving along n in the original frame, because the Euler flux
//    is rotation covariant: with u_n = v.n,
//        F(Q; n) = (rho u_n, u_n (g2 p + ke), rho u_n x, rho u_n a, rho v u_n + p n),
//    and the HLLC star state is A (1, E/rho + (S_c - u_n)(S_c + p/(rho (S - u_n))),
//    x, a, v + (S_c - u_n) n).  For the unit vector n1 (velocity jump,
//    oriented so that n1_d >= 0) and face axis d, the second normal is
//    n2 = (n1 x d) x n1 / |.| = (d - n1_d n1) / s with s = |n1 - n1_d d|,
//    so n2_d = s and the hybrid weights |d.n1|, |d.n2| are n1_d and s.
//  * both solves share the Roe-average factors, which depend only on the
//    two densities and sound speeds.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// FP32 slope limiters of the mixed path.  The mixed kernel works on a state
// scaled by powers of two (largest density and pressure near 2^16), so the
// squares below stay in FP32 range; the reference's a b <= 1e-30 test is
// applied with the threshold scaled alike (thr).
__device__ __forceinline__ float albada_f(float a, float b, float thr) {
    const float ab = a * b;
    return ab <= thr ? 0.f : ab * (a + b) * frcp(a * a + b * b);
}

__device__ __forceinline__ float limiter_f(int lim, float a, float b, float thr) {
    if (lim == LIM_ALBADA) return albada_f(a, b, thr);
    if (lim == LIM_OSPRE) {
        if (a * b <= thr) return 0.f;
        return 1.5f * b * (a * a + a * b) / (a * a + a * b + b * b);
    }
    return limiter<float>(lim, a, b);  // minmod, maxmod, MC, superbee, van Leer: no threshold
}

struct SideF {
    float rho, p, x, a, v[3];
    float sr;   // sqrt(rho)
    float a2;   // gamma p / rho
    float ke;   // rho |v|^2 / 2
    float E;    // p / (gamma - 1) + ke
};

__device__ __forceinline__ SideF side_f(const float q[7], float g, float igm1) {
    SideF s;
    s.rho = q[DEN];
    s.p = q[PRE];
    s.x = q[HII];
    s.a = q[ADV];
    s.v[0] = q[VEL + 0];
    s.v[1] = q[VEL + 1];
    s.v[2] = q[VEL + 2];
    // one reciprocal square root gives both sqrt(rho) and 1/rho
    const float rs = __builtin_amdgcn_rsqf(s.rho);
    s.sr = s.rho * rs;
    s.a2 = g * s.p * (rs * rs);
    s.ke = 0.5f * s.rho * (s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]);
    s.E = s.p * igm1 + s.ke;
    return s;
}

__device__ __forceinline__ float dotn(const float v[3], const float n[3]) { return v[0] * n[0] + v[1] * n[1] + v[2] * n[2]; }

// F(Q; n) for a side with normal velocity un.
__device__ __forceinline__ void flux_n(float F[7], const SideF& s, float un, const float n[3], float g2) {
    const float m = s.rho * un;
    F[DEN] = m;
    F[PRE] = un * (g2 * s.p + s.ke);
    F[HII] = m * s.x;
    F[ADV] = m * s.a;
#pragma unroll
    for (int i = 0; i < 3; ++i) F[VEL + i] = m * s.v[i] + s.p * n[i];
}

__device__ __forceinline__ void cons_of(float U[7], const SideF& s) {
    U[DEN] = s.rho;
    U[PRE] = s.E;
    U[HII] = s.rho * s.x;
    U[ADV] = s.rho * s.a;
#pragma unroll
    for (int i = 0; i < 3; ++i) U[VEL + i] = s.rho * s.v[i];
}

// Roe-average pieces shared by every normal: 1/(sr_l + sr_r), the mixed
// sound speed term and the velocity-jump weight nu2.
struct RoeF {
    float is, base2, nu2;
};

__device__ __forceinline__ RoeF roe_f(const SideF& L, const SideF& R) {
    RoeF r;
    r.is = frcp(L.sr + R.sr);
    r.base2 = (L.sr * L.a2 + R.sr * R.a2) * r.is;
    r.nu2 = 0.5f * L.sr * R.sr * r.is * r.is;
    return r;
}

__device__ __forceinline__ void speeds_f(const SideF& L, const SideF& R, const RoeF& r, float unl, float unr,
                                         float& S_l, float& S_r) {
    const float ut = (L.sr * unl + R.sr * unr) * r.is;
    const float du = unr - unl;
    const float d = fsqrt(r.base2 + r.nu2 * du * du);
    S_l = ut - d;
    S_r = ut + d;
}

// HLL flux along unit normal n (HartenLaxLeerSolver::solve).
__device__ __forceinline__ void hll_f(float F[7], const SideF& L, const SideF& R, const RoeF& r, const float n[3],
                                      float g2) {
    const float unl = dotn(L.v, n), unr = dotn(R.v, n);
    float S_l, S_r;
    speeds_f(L, R, r, unl, unr, S_l, S_r);
    float Fl[7], Fr[7];
    flux_n(Fl, L, unl, n, g2);
    flux_n(Fr, R, unr, n, g2);
    if (S_r <= 0.f) {
#pragma unroll
        for (int i = 0; i < 7; ++i) F[i] = Fr[i];
    } else if (S_l >= 0.f) {
#pragma unroll
        for (int i = 0; i < 7; ++i) F[i] = Fl[i];
    } else {
        float Ul[7], Ur[7];
        cons_of(Ul, L);
        cons_of(Ur, R);
        // weights S_r / (S_r - S_l) in (0, 1) and S_l / (S_r - S_l) in
        // (-1, 0) first: no product is larger than a flux or a speed times a
        // conserved jump (S_l S_r dE alone overflows FP32 for energies of
        // 1e22 code units)
        const float iw = frcp(S_r - S_l), wr = S_r * iw, wl = S_l * iw, slr = S_l * wr;
#pragma unroll
        for (int i = 0; i < 7; ++i) F[i] = wr * Fl[i] - wl * Fr[i] + slr * (Ur[i] - Ul[i]);
    }
}

// HLLC flux along unit normal n (HartenLaxLeerContactSolver::solve).
__device__ __forceinline__ void hllc_f(float F[7], const SideF& L, const SideF& R, const RoeF& r, const float n[3],
                                       float g2) {
    const float unl = dotn(L.v, n), unr = dotn(R.v, n);
    float S_l, S_r;
    speeds_f(L, R, r, unl, unr, S_l, S_r);
    const float wl = L.rho * (S_l - unl), wr = R.rho * (S_r - unr);
    const float S_c = (R.p - L.p + wl * unl - wr * unr) * frcp(wl - wr);
    // supersonic cases take the upwind flux unchanged
    const bool sub = !(S_l >= 0.f) && !(S_r <= 0.f);
    const bool left = (S_l >= 0.f) || (sub && S_c >= 0.f);
    const SideF& K = left ? L : R;
    const float un = left ? unl : unr;
    const float S = left ? S_l : S_r;
    flux_n(F, K, un, n, g2);
    if (sub) {
        // A = rho w / (S - S_c); A E / rho = w E / (S - S_c) and
        // A p / (rho w) = p / (S - S_c): one reciprocal for the star state
        const float w = S - un;
        const float iss = frcp(S - S_c);
        const float A = K.rho * w * iss;
        const float dsc = S_c - un;
        float dU[7];
        dU[DEN] = A - K.rho;
        dU[PRE] = w * iss * K.E + dsc * (A * S_c + K.p * iss) - K.E;
        dU[HII] = (A - K.rho) * K.x;
        dU[ADV] = (A - K.rho) * K.a;
#pragma unroll
        for (int i = 0; i < 3; ++i) dU[VEL + i] = A * (K.v[i] + dsc * n[i]) - K.rho * K.v[i];
#pragma unroll
        for (int i = 0; i < 7; ++i) F[i] += S * dU[i];
    }
}

// Mixed-precision face flux for solver SOLVER along axis DIM.  Returns true
// if the HII flux is NaN.
template <int SOLVER, int DIM>
__device__ __forceinline__ bool riemann_f(float F[7], const float* Ql, const float* Qr, const GasConsts<float>& c,
                                          float igm1, float nn_thr) {
    const SideF L = side_f(Ql, c.g, igm1), R = side_f(Qr, c.g, igm1);
    const RoeF r = roe_f(L, R);
    const float d[3] = {DIM == 0 ? 1.f : 0.f, DIM == 1 ? 1.f : 0.f, DIM == 2 ? 1.f : 0.f};
    if constexpr (SOLVER == SOLVER_HLL) {
        hll_f(F, L, R, r, d, c.g2);
    } else if constexpr (SOLVER == SOLVER_HLLC) {
        hllc_f(F, L, R, r, d, c.g2);
    } else {
        constexpr int A = DIM == 0 ? 1 : 0, B = DIM == 2 ? 1 : 2;  // transverse axes
        float n1[3] = {R.v[0] - L.v[0], R.v[1] - L.v[1], R.v[2] - L.v[2]};
        if (n1[DIM] < 0.f) {
#pragma unroll
            for (int i = 0; i < 3; ++i) n1[i] = -n1[i];
        }
        const float nn = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
        if (nn < nn_thr) {  // |n1| < 1e-6 (code units): pure HLLC along the face axis
            hllc_f(F, L, R, r, d, c.g2);
        } else {
            const float in = __builtin_amdgcn_rsqf(nn);
#pragma unroll
            for (int i = 0; i < 3; ++i) n1[i] *= in;
            const float s2 = n1[A] * n1[A] + n1[B] * n1[B];
            if (s2 == 0.f) {  // n1 along the face axis: pure HLL
                hll_f(F, L, R, r, d, c.g2);
            } else {
                const float is = __builtin_amdgcn_rsqf(s2), s = s2 * is;
                float n2[3];
                n2[DIM] = s;
                n2[A] = -n1[A] * n1[DIM] * is;
                n2[B] = -n1[B] * n1[DIM] * is;
                float F1[7], F2[7];
                hll_f(F1, L, R, r, n1, c.g2);
                hllc_f(F2, L, R, r, n2, c.g2);
                const float w1 = n1[DIM];
#pragma unroll
                for (int i = 0; i < 7; ++i) F[i] = w1 * F1[i] + s * F2[i];
            }
        }
    }
    return F[HII] != F[HII];
}

}  // namespace tg
