// hydro.hip -- fused hydro stage kernels for CDNA4 (gfx950).
//
// One launch = one stage of the reference's midpoint predictor-corrector
// (Torch::hydroStep, src/Torch/Torch.cpp:366-387) over the whole grid:
//
//   Q   = QfromU(Usrc)                       (+ fixPrimitives in stage 1)
//   ghost cells by boundary condition        Grid::applyBCs, Grid.cpp:649-802
//   QL/QR = linear reconstruction            Hydrodynamics::reconstruct, Hydro.cpp:44-67
//   F    = Riemann solve per face            Hydrodynamics::calcFluxes, Hydro.cpp:91-115
//   UDOT = sum_dim (A/V)(F_left - F_right)   Hydrodynamics::updateSourceTerms, Hydro.cpp:117-137
//   Udst = fixSolution(Ubase + dtf * UDOT)   Fluid::advSolution/fixSolution, Fluid.cpp:80-136
//   stage 2 only: max_cells inv_t of fixPrimitives(QfromU(Udst)), the CFL
//   reduction of the next step (Hydro.cpp:77-89), via wave shuffles + one
//   atomicMax per block on the bit pattern of a non-negative double.
//
// Data layout: conserved state U is structure-of-arrays in HBM,
// U[v * ncell + x + nx * (y + ny * z)], v in UID order, x fastest.
// A workgroup owns an 8x8x8 tile of cells (576 threads = 9 waves).  Q of the
// tile lives in LDS; for each sweep direction the 2-deep halo slabs on both
// sides are staged in LDS, every one of the 9x8x8 faces of that direction is
// solved once by one thread, the fluxes go to LDS, and each cell thread
// accumulates its two faces in the reference's order.  Tiles are assigned
// XCD-contiguously (consecutive workgroups round-robin over the 8 XCDs, so
// workgroup b is given tile (b % 8) * (ntiles / 8) + b / 8) so that each
// XCD's L2 sees neighbouring tiles' halos.
//
// For precision == 1 (mixed) Q is rounded to float after the FP64 pressure
// recovery, reconstruction and the Riemann solve run in float, and the
// fluxes are accumulated, applied and floored in FP64.

#include "hydro_math.hpp"
#include "torchgpu_internal.hpp"

namespace tg {

constexpr int TILE = 8;
constexpr int TCELLS = TILE * TILE * TILE;        // 512
constexpr int TFACES = (TILE + 1) * TILE * TILE;  // 576
constexpr int THALO = 4 * TILE * TILE;            // 256 (2 layers x 2 sides)
constexpr int NTHREADS = TFACES;

__device__ __forceinline__ int tile_of_block(int b, int ntiles) {
    const int q = ntiles >> 3, r = ntiles & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

// fixPrimitives for one cell (Fluid.cpp:138-150).  u_den is U[DEN] of the
// same cell, used by the reference's temperature floor.
__device__ __forceinline__ void fix_primitives(const HydroGeom& g, double q[7], double u_den) {
    q[HII] = smax(smin(q[HII], 1.0), 0.0);
    q[ADV] = smax(smin(q[ADV], 1.0), 0.0);
    q[DEN] = smax(q[DEN], g.dfloor);
    q[PRE] = smax(q[PRE], g.pfloor);
    double mu_inv = g.X * (q[HII] + 1.0) + g.mu_c;
    double temperature = q[PRE] / (mu_inv * g.Rgas * q[DEN]);
    if (temperature < g.tfloor) q[PRE] = mu_inv * g.Rgas * u_den * g.tfloor;
}

// fixSolution for one cell (Fluid.cpp:89-136).  Returns true if the
// reference would throw.
__device__ __forceinline__ bool fix_solution(const HydroGeom& g, double u[7]) {
    bool bad = !isfinite(u[DEN]) || !isfinite(u[PRE]);
    double hii = smax(smin(u[HII] / u[DEN], 1.0), 0.0);
    double adv = smax(smin(u[ADV] / u[DEN], 1.0), 0.0);
    double den = smax(u[DEN], g.dfloor);
    double v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = u[VEL + d] / u[DEN];
    double ke = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) ke += v[d] * v[d];
    ke *= 0.5 * u[DEN];
    double pre = (u[PRE] - ke) * g.gm1;
    ke *= den / u[DEN];
    if (pre < g.pfloor) pre = g.pfloor;
    double mu_inv = g.X * (hii + 1.0) + g.mu_c;
    double temperature = pre / (mu_inv * g.Rgas * den);
    if (temperature < g.tfloor) pre = mu_inv * g.Rgas * den * g.tfloor;
    u[DEN] = den;
    u[PRE] = pre / g.gm1 + ke;
    u[HII] = hii * den;
    u[ADV] = adv * den;
#pragma unroll
    for (int d = 0; d < 3; ++d) u[VEL + d] = den * v[d];
    bad |= (u[DEN] == 0.0 || u[PRE] == 0.0);
#pragma unroll
    for (int i = 0; i < 7; ++i) bad |= (u[i] != u[i]) || isinf(u[i]);
    return bad;
}

// CFL term of one cell, Hydro.cpp:80-85 (inv_t == 0 -> 1/dt_max).
__device__ __forceinline__ double cfl_inv_t(const HydroGeom& g, const double q[7]) {
    double inv_t = 0;
    double ss = sqrt(g.g * q[PRE] / q[DEN]);
#pragma unroll
    for (int d = 0; d < 3; ++d) inv_t += (fabs(q[VEL + d]) + ss) / g.dx;
    if (inv_t == 0) inv_t = 1.0 / g.dt_max;
    return inv_t;
}

// Mixed-precision path: the same FP64 floors, primitive recovery and CFL
// term, with each run of divisions by one quantity replaced by one
// reciprocal and the temperature tests written as products.
__device__ __forceinline__ void q_from_u_mixed(double q[7], const double u[7], double gm1) {
    const double ir = 1.0 / u[DEN];
    q[DEN] = u[DEN];
#pragma unroll
    for (int d = 0; d < 3; ++d) q[VEL + d] = u[VEL + d] * ir;
    const double ke = 0.5 * (u[VEL + 0] * q[VEL + 0] + u[VEL + 1] * q[VEL + 1] + u[VEL + 2] * q[VEL + 2]);
    q[PRE] = (u[PRE] - ke) * gm1;
    q[HII] = u[HII] * ir;
    q[ADV] = u[ADV] * ir;
}

__device__ __forceinline__ void fix_primitives_mixed(const HydroGeom& g, double q[7], double u_den) {
    q[HII] = smax(smin(q[HII], 1.0), 0.0);
    q[ADV] = smax(smin(q[ADV], 1.0), 0.0);
    q[DEN] = smax(q[DEN], g.dfloor);
    q[PRE] = smax(q[PRE], g.pfloor);
    const double muR = (g.X * (q[HII] + 1.0) + g.mu_c) * g.Rgas;
    if (q[PRE] < g.tfloor * muR * q[DEN]) q[PRE] = muR * u_den * g.tfloor;
}

__device__ __forceinline__ bool fix_solution_mixed(const HydroGeom& g, double u[7]) {
    bool bad = !isfinite(u[DEN]) || !isfinite(u[PRE]);
    const double ir = 1.0 / u[DEN];
    const double hii = smax(smin(u[HII] * ir, 1.0), 0.0);
    const double adv = smax(smin(u[ADV] * ir, 1.0), 0.0);
    const double den = smax(u[DEN], g.dfloor);
    double v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = u[VEL + d] * ir;
    double ke = 0.5 * u[DEN] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double pre = (u[PRE] - ke) * g.gm1;
    ke *= den * ir;
    if (pre < g.pfloor) pre = g.pfloor;
    const double muR = (g.X * (hii + 1.0) + g.mu_c) * g.Rgas;
    if (pre < g.tfloor * muR * den) pre = muR * den * g.tfloor;
    u[DEN] = den;
    u[PRE] = pre * g.igm1 + ke;
    u[HII] = hii * den;
    u[ADV] = adv * den;
#pragma unroll
    for (int d = 0; d < 3; ++d) u[VEL + d] = den * v[d];
    bad |= (u[DEN] == 0.0 || u[PRE] == 0.0);
#pragma unroll
    for (int i = 0; i < 7; ++i) bad |= (u[i] != u[i]) || isinf(u[i]);
    return bad;
}

__device__ __forceinline__ double cfl_inv_t_mixed(const HydroGeom& g, const double q[7]) {
    const double ss = sqrt(g.g * q[PRE] / q[DEN]);
    double inv_t = (fabs(q[VEL + 0]) + ss + (fabs(q[VEL + 1]) + ss) + (fabs(q[VEL + 2]) + ss)) * g.idx;
    if (inv_t == 0) inv_t = 1.0 / g.dt_max;
    return inv_t;
}

// Records an error of the current step unless an earlier step already
// failed (the reference stops at its first exception).
__device__ __forceinline__ void set_error(StepState* st, unsigned int bit) {
    const int s1 = st->step + 1;
    const int prev = atomicCAS(&st->err_step1, 0, s1);
    if (prev == 0 || prev == s1) atomicOr(&st->err, bit);
}

// Block-wide max of non-negative doubles (NaN entries ignored) -> one
// atomicMax on the bit pattern.  All NTHREADS threads must call it.
__device__ __forceinline__ void block_max_to(double v, unsigned long long* dst, double* red) {
    if (!(v >= 0.0)) v = 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < NTHREADS / 64; ++w) m = fmax(m, red[w]);
        atomicMax(dst, (unsigned long long)__double_as_longlong(m));
    }
}

// Source interior coordinate and velocity sign for a ghost position along
// one dimension (Grid.cpp:661-802).  pos is the global coordinate, -2..N+1.
// Returns the interior index to read; flip = -1 negates the normal velocity,
// +1 keeps it, 0 = "flip if the interior velocity points inwards" (outflow)
// or outwards (inflow) -- resolved by the caller.
struct GhostMap {
    int src;
    int mode;  // 0 copy, 1 negate, 2 outflow-left, 3 outflow-right, 4 inflow-left, 5 inflow-right
};

__device__ __forceinline__ GhostMap ghost_map(int pos, int N, int bcl, int bcr) {
    GhostMap m{pos, 0};
    if (pos < 0) {
        const int L = -pos;  // layer 1 or 2
        switch (bcl) {
        case BC_PERIODIC:   m.src = N - L; break;
        case BC_REFLECTING: m.src = L; m.mode = 1; break;
        case BC_OUTFLOW:    m.src = 0; m.mode = 2; break;
        case BC_INFLOW:     m.src = 0; m.mode = 4; break;
        default:            m.src = 0; break;
        }
    } else if (pos >= N) {
        const int L = pos - N + 1;
        switch (bcr) {
        case BC_PERIODIC:   m.src = L - 1; break;
        case BC_REFLECTING: m.src = N - 1 - L; m.mode = 1; break;
        case BC_OUTFLOW:    m.src = N - 1; m.mode = 3; break;
        case BC_INFLOW:     m.src = N - 1; m.mode = 5; break;
        default:            m.src = N - 1; break;
        }
    }
    // keep every read in bounds (positions far outside the domain are never
    // used by a valid face)
    m.src = m.src < 0 ? 0 : (m.src >= N ? N - 1 : m.src);
    return m;
}

__device__ __forceinline__ void flip_normal(double& vn, int mode) {
    switch (mode) {
    case 1: vn = -vn; break;
    case 2: if (vn > 0) vn = -1.0 * vn; break;
    case 3: if (vn < 0) vn = -1.0 * vn; break;
    case 4: if (vn < 0) vn = -1.0 * vn; break;
    case 5: if (vn > 0) vn = -1.0 * vn; break;
    default: break;
    }
}

__device__ __forceinline__ void apply_ghost_mode(double q[7], int mode, int dim) {
    // static register indices only (no dynamically indexed private array)
    if (dim == 0) flip_normal(q[VEL + 0], mode);
    else if (dim == 1) flip_normal(q[VEL + 1], mode);
    else flip_normal(q[VEL + 2], mode);
}

// Load Q of the cell at global (x, y, z), which may lie in the 2-deep ghost
// layer of one dimension (the first out-of-range dimension decides).
template <int STAGE, bool MIXED>
__device__ __forceinline__ void load_q(const HydroGeom& g, const double* __restrict__ U, int x, int y, int z,
                                       double q[7], StepState* st, bool check) {
    int c[3] = {x, y, z};
    const int N[3] = {g.nx, g.ny, g.nz};
    int mode = 0, mdim = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if ((c[d] < 0 || c[d] >= N[d]) && mode == 0 && mdim == 0) {
            GhostMap m = ghost_map(c[d], N[d], g.bc_l[d], g.bc_r[d]);
            c[d] = m.src;
            mode = m.mode;
            mdim = d + 1;
        } else {
            c[d] = c[d] < 0 ? 0 : (c[d] >= N[d] ? N[d] - 1 : c[d]);
        }
    }
    const long long idx = (long long)c[0] + (long long)g.nx * ((long long)c[1] + (long long)g.ny * c[2]);
    double u[7];
#pragma unroll
    for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + idx];
    if constexpr (MIXED) q_from_u_mixed(q, u, g.gm1);
    else QfromU(q, u, g.gm1);
    if (STAGE == 1) {
        if constexpr (MIXED) fix_primitives_mixed(g, q, u[DEN]);
        else fix_primitives(g, q, u[DEN]);
        if (check) {
            // Torch::checkValues("hydro before"), Torch.cpp:420-443
            bool bad = (q[DEN] == 0.0) || (q[PRE] == 0.0);
#pragma unroll
            for (int v = 0; v < 7; ++v) bad |= (u[v] != u[v]) || isinf(u[v]);
            if (bad) set_error(st, ERR_CHECK_VALUES);
        }
    }
    if (mdim) apply_ghost_mode(q, mode, mdim - 1);
}

template <typename R>
struct TileLDS {
    R qc[7][TCELLS];
    R qh[7][THALO];
    R fl[7][TFACES];
    double red[NTHREADS / 64];
};

// LDS address of the Q value at along-position pos (-2..9) of sweep DIM with
// transverse coordinates (ta, tb).
template <int DIM>
__device__ __forceinline__ int center_index(int pos, int ta, int tb) {
    if (DIM == 0) return pos + TILE * (ta + TILE * tb);
    if (DIM == 1) return ta + TILE * (pos + TILE * tb);
    return ta + TILE * (tb + TILE * pos);
}

template <typename R, int DIM>
__device__ __forceinline__ void q_at(const TileLDS<R>& s, int pos, int ta, int tb, R q[7]) {
    if (pos >= 0 && pos < TILE) {
        const int i = center_index<DIM>(pos, ta, tb);
#pragma unroll
        for (int v = 0; v < 7; ++v) q[v] = s.qc[v][i];
    } else {
        const int h = pos < 0 ? pos + 2 : pos - TILE + 2;
        const int i = h * (TILE * TILE) + tb * TILE + ta;
#pragma unroll
        for (int v = 0; v < 7; ++v) q[v] = s.qh[v][i];
    }
}

// Flux through face f (between along-positions f-1 and f) of pencil (ta, tb)
// of sweep DIM: reconstruction from the tile's LDS Q and the Riemann solve.
// Returns true if the HII flux is NaN.
template <int SOLVER, typename R, int DIM>
__device__ __forceinline__ bool face_flux(const HydroGeom& g, const TileLDS<R>& s, int f, int ta, int tb,
                                          const GasConsts<R>& gc, R F[7]) {
    R ql[7], qr[7];
    if (g.spatial_order == 1) {
        R a[7], b[7], cc[7], dd[7];
        q_at<R, DIM>(s, f - 2, ta, tb, a);
        q_at<R, DIM>(s, f - 1, ta, tb, b);
        q_at<R, DIM>(s, f, ta, tb, cc);
        q_at<R, DIM>(s, f + 1, ta, tb, dd);
        bool fast = false;
        if constexpr (sizeof(R) == 4) {
            if (g.limiter == LIM_ALBADA) {
                fast = true;
#pragma unroll
                for (int v = 0; v < 7; ++v) {
                    ql[v] = b[v] + 0.5f * albada_f(b[v] - a[v], cc[v] - b[v]);
                    qr[v] = cc[v] - 0.5f * albada_f(cc[v] - b[v], dd[v] - cc[v]);
                }
            }
        }
        if (!fast) {
#pragma unroll
            for (int v = 0; v < 7; ++v) {
                R dl = limiter<R>(g.limiter, b[v] - a[v], cc[v] - b[v]);
                ql[v] = b[v] + R(0.5) * dl;   // QR of cell f-1
                R dr = limiter<R>(g.limiter, cc[v] - b[v], dd[v] - cc[v]);
                qr[v] = cc[v] - R(0.5) * dr;  // QL of cell f
            }
        }
    } else {
        q_at<R, DIM>(s, f - 1, ta, tb, ql);
        q_at<R, DIM>(s, f, ta, tb, qr);
    }
    if constexpr (sizeof(R) == 4) {
        // mixed precision: hardware reciprocals, rotation-free hybrid flux
        return riemann_f<SOLVER, DIM>(F, ql, qr, gc, g.igm1f);
    } else {
        R a_l2 = gc.g * ql[PRE] / ql[DEN];
        R a_r2 = gc.g * qr[PRE] / qr[DEN];
        return riemann<SOLVER, R, DIM>(F, ql, qr, a_l2, a_r2, gc);
    }
}

template <int SOLVER, typename R, int STAGE, int DIM>
__device__ __forceinline__ void sweep(const HydroGeom& g, TileLDS<R>& s, const double* __restrict__ Usrc,
                                      const int o[3], double udot[7], StepState* st, const GasConsts<R>& gc) {
    const int tid = threadIdx.x;
    const int N[3] = {g.nx, g.ny, g.nz};
    constexpr int DA = DIM == 0 ? 1 : 0;       // transverse dim "a"
    constexpr int DB = DIM == 2 ? 1 : 2;       // transverse dim "b"

    // 1. halo slabs of this direction
    if (tid < THALO) {
        const int h = tid >> 6, ta = tid & 7, tb = (tid >> 3) & 7;
        const int pos = h < 2 ? h - 2 : TILE + h - 2;
        int c[3];
        c[DIM] = o[DIM] + pos;
        c[DA] = o[DA] + ta;
        c[DB] = o[DB] + tb;
        double q[7];
        load_q<STAGE, (sizeof(R) == 4)>(g, Usrc, c[0], c[1], c[2], q, st, false);
#pragma unroll
        for (int v = 0; v < 7; ++v) s.qh[v][tid] = (R)q[v];
    }
    __syncthreads();

    // 2. one face per thread: face f sits between along-positions f-1 and f
    {
        const int f = tid % (TILE + 1), rest = tid / (TILE + 1);
        const int ta = rest & 7, tb = rest >> 3;
        const bool valid = (o[DIM] + f <= N[DIM]) && (o[DA] + ta < N[DA]) && (o[DB] + tb < N[DB]);
        R F[7];
        bool bad = false;
        if (valid) bad = face_flux<SOLVER, R, DIM>(g, s, f, ta, tb, gc, F);
        else {
#pragma unroll
            for (int v = 0; v < 7; ++v) F[v] = R(0);
        }
        if (bad) set_error(st, STAGE == 1 ? ERR_RIEMANN_NAN_1 : ERR_RIEMANN_NAN_2);
#pragma unroll
        for (int v = 0; v < 7; ++v) s.fl[v][tid] = F[v];
    }
    __syncthreads();

    // 3. flux divergence in the reference order: += (A/V) F_L, -= (A/V) F_R
    if (tid < TCELLS) {
        const int lc[3] = {tid & 7, (tid >> 3) & 7, tid >> 6};
        const int a = lc[DIM], ta = lc[DA], tb = lc[DB];
        const int fL = (tb * TILE + ta) * (TILE + 1) + a;
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            double u = udot[v];
            u += g.kflux * (double)s.fl[v][fL];
            u -= g.kflux * (double)s.fl[v][fL + 1];
            udot[v] = u;
        }
    }
}

// dtf_host >= 0 overrides the stage's dt (st->dt / 2.0 for the predictor,
// st->dt for the corrector) with a host-given, already scaled dt.
template <int SOLVER, typename R, int STAGE>
__global__ __launch_bounds__(NTHREADS) void hydro_stage_kernel(HydroGeom g, const double* __restrict__ Usrc,
                                                              const double* Ubase, double* Udst, StepState* st,
                                                              double dtf_host) {
    __shared__ TileLDS<R> s;
    if (st->halted) return;
    const int tile = tile_of_block(blockIdx.x, g.ntiles);
    const int tx = tile % g.ntx, ty = (tile / g.ntx) % g.nty, tz = tile / (g.ntx * g.nty);
    const int o[3] = {tx * TILE, ty * TILE, tz * TILE};
    const int tid = threadIdx.x;
    const GasConsts<R> gc = {(R)(sizeof(R) == 8 ? g.g : g.gf), (R)(sizeof(R) == 8 ? g.gm1 : g.gm1f),
                             (R)(sizeof(R) == 8 ? g.g2 : g.g2f)};

    // tile cells: Q into LDS (ghost mapping for positions outside the grid)
    const int lx = tid & 7, ly = (tid >> 3) & 7, lz = tid >> 6;
    const int gx = o[0] + lx, gy = o[1] + ly, gz = o[2] + lz;
    const bool active = tid < TCELLS && gx < g.nx && gy < g.ny && gz < g.nz;
    if (tid < TCELLS) {
        double q[7];
        load_q<STAGE, (sizeof(R) == 4)>(g, Usrc, gx, gy, gz, q, st, active);
#pragma unroll
        for (int v = 0; v < 7; ++v) s.qc[v][tid] = (R)q[v];
    }
    double udot[7] = {0, 0, 0, 0, 0, 0, 0};

    sweep<SOLVER, R, STAGE, 0>(g, s, Usrc, o, udot, st, gc);
    sweep<SOLVER, R, STAGE, 1>(g, s, Usrc, o, udot, st, gc);
    sweep<SOLVER, R, STAGE, 2>(g, s, Usrc, o, udot, st, gc);

    double inv_t = 0.0;
    if (active) {
        // gravity source with GRAV = 0 (Hydro.cpp:134-137)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double grav = 0.0;
            udot[VEL + i] += grav;
            udot[PRE] += (double)s.qc[VEL + i][tid] * grav;
        }
        const long long c = (long long)gx + (long long)g.nx * ((long long)gy + (long long)g.ny * gz);
        const double dtf = dtf_host >= 0.0 ? dtf_host : (STAGE == 1 ? st->dt / 2.0 : st->dt);
        double u[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = Ubase[v * g.ncell + c];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] += dtf * udot[v];
        const bool bad = sizeof(R) == 4 ? fix_solution_mixed(g, u) : fix_solution(g, u);
        if (bad) set_error(st, STAGE == 1 ? ERR_FIX_SOLUTION_1 : ERR_FIX_SOLUTION_2);
#pragma unroll
        for (int v = 0; v < 7; ++v) Udst[v * g.ncell + c] = u[v];
        if (STAGE == 2) {
            double q[7];
            if constexpr (sizeof(R) == 4) {
                q_from_u_mixed(q, u, g.gm1);
                fix_primitives_mixed(g, q, u[DEN]);
                inv_t = cfl_inv_t_mixed(g, q);
            } else {
                QfromU(q, u, g.gm1);
                fix_primitives(g, q, u[DEN]);
                inv_t = cfl_inv_t(g, q);
            }
        }
    }
    if (STAGE == 2) block_max_to(inv_t, &st->maxinv, s.red);
}

// Standalone CFL reduction over U (Hydrodynamics::calculateTimeStep on
// fixPrimitives(QfromU(U))).
__global__ __launch_bounds__(256) void cfl_kernel(HydroGeom g, const double* __restrict__ U, StepState* st) {
    __shared__ double red[4];
    double m = 0.0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long long)gridDim.x * blockDim.x) {
        double u[7], q[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + c];
        QfromU(q, u, g.gm1);
        fix_primitives(g, q, u[DEN]);
        double it = cfl_inv_t(g, q);
        if (it >= 0.0) m = fmax(m, it);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(&st->maxinv, (unsigned long long)__double_as_longlong(b));
    }
}

// Start of a step: dt = min(dt_nextCheckpoint, dt_cfl) as Torch::run +
// Torch::fullStep + Torch::calculateTimeStep (Torch.cpp:266-270,304-339,397),
// with the checkpoint arithmetic of Checkpointer::update (Checkpointer.cpp:14-26).
// use_cfl = 0 takes dt_override instead of the CFL/first-step value.
__global__ void step_begin_kernel(StepState* st, int use_cfl, double dt_override) {
    if (!use_cfl) {  // an explicit Torch::hydroStep(dt): no clock rules apply
        st->halted = 0;
        st->dt = dt_override;
        return;
    }
    if (st->halted) return;
    if (!(st->time < st->tmax)) {
        st->halted = 1;
        return;
    }
    const double delta = 100.0 / st->ncheckpoints;
    double pct = smin(100.0, 100.0 * st->time / st->tmax);
    int count = st->ckpt_count;
    while (pct - count * delta > delta - 1.0e-8) count += 1;
    st->ckpt_count = count;
    const double dt_next = ((count + 1) * delta - pct) * st->tmax / 100.0;
    double dt;
    if (st->first) {
        dt = st->dt_max * 1.0e-20;
        st->first = 0;
    } else {
        const double maxinv = __longlong_as_double((long long)st->maxinv);
        double tmin = st->dt_max;
        if (maxinv > 0.0) tmin = smin(tmin, 0.5 / maxinv);
        dt = smin(tmin, st->dt_max);
    }
    st->dt = smin(dt_next, dt);
    st->maxinv = 0ull;
}

__global__ void step_end_kernel(StepState* st) {
    if (st->halted) return;
    st->time += st->dt;
    st->step += 1;
}

__global__ void q_from_u_kernel(HydroGeom g, const double* __restrict__ U, double* __restrict__ Q, int fix) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long long)gridDim.x * blockDim.x) {
        double u[7], q[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + c];
        QfromU(q, u, g.gm1);
        if (fix) fix_primitives(g, q, u[DEN]);
#pragma unroll
        for (int v = 0; v < 7; ++v) Q[v * g.ncell + c] = q[v];
    }
}

// Torch::initialise tail (Torch.cpp:117-119): fixPrimitives on the uploaded
// Q (U is still zero there, so the temperature floor reads U[DEN] = 0), then
// globalUfromQ.  Q is fixed in place.
__global__ void u_from_q_kernel(HydroGeom g, double* __restrict__ Q, double* __restrict__ U) {
    const GasConsts<double> gc = {g.g, g.gm1, g.g2};
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long long)gridDim.x * blockDim.x) {
        double u[7], q[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) q[v] = Q[v * g.ncell + c];
        fix_primitives(g, q, 0.0);
        UfromQ<double>(u, q, gc);
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            Q[v * g.ncell + c] = q[v];
            U[v * g.ncell + c] = u[v];
        }
    }
}

// Thermodynamics::updateSourceTerms + Fluid::advSolution + fixSolution on
// the grid: UDOT[PRE] = rate (other UDOT zero), U += dtf * UDOT, then
// fixSolution (Thermodynamics.cpp:478-485, Fluid.cpp:80-136).
__global__ void apply_rate_fix_kernel(HydroGeom g, double* __restrict__ U, const double* __restrict__ rate, double dtf,
                                      StepState* st) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long long)gridDim.x * blockDim.x) {
        double u[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + c];
        double udot[7] = {0, 0, 0, 0, 0, 0, 0};
        udot[PRE] += rate ? rate[c] : 0.0;
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] += dtf * udot[v];
        if (fix_solution(g, u)) set_error(st, ERR_THERMO_FIX);
#pragma unroll
        for (int v = 0; v < 7; ++v) U[v * g.ncell + c] = u[v];
    }
}

hipError_t launch_fix_solution(const HydroGeom& g, double* U, const double* rate, double dtf, StepState* st,
                               hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(apply_rate_fix_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, U, rate, dtf, st);
    return hipGetLastError();
}

template <int SOLVER, typename R>
static hipError_t launch_stage_t(const HydroGeom& g, int stage, const double* Usrc, const double* Ubase, double* Udst,
                                 StepState* st, double dtf, hipStream_t stream) {
    dim3 grid(g.ntiles), block(NTHREADS);
    if (stage == 1)
        hipLaunchKernelGGL((hydro_stage_kernel<SOLVER, R, 1>), grid, block, 0, stream, g, Usrc, Ubase, Udst, st, dtf);
    else
        hipLaunchKernelGGL((hydro_stage_kernel<SOLVER, R, 2>), grid, block, 0, stream, g, Usrc, Ubase, Udst, st, dtf);
    return hipGetLastError();
}

hipError_t launch_hydro_stage(const HydroGeom& g, int solver, int precision, int stage, const double* Usrc,
                              const double* Ubase, double* Udst, StepState* st, hipStream_t stream, double dtf) {
    if (precision == 0) {
        switch (solver) {
        case SOLVER_HLL:  return launch_stage_t<SOLVER_HLL, double>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
        case SOLVER_HLLC: return launch_stage_t<SOLVER_HLLC, double>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
        default:          return launch_stage_t<SOLVER_ROTATED, double>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
        }
    }
    switch (solver) {
    case SOLVER_HLL:  return launch_stage_t<SOLVER_HLL, float>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    case SOLVER_HLLC: return launch_stage_t<SOLVER_HLLC, float>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    default:          return launch_stage_t<SOLVER_ROTATED, float>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    }
}

hipError_t launch_cfl(const HydroGeom& g, const double* U, StepState* st, hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(cfl_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, U, st);
    return hipGetLastError();
}

hipError_t launch_step_begin(StepState* st, int use_cfl, double dt_override, hipStream_t stream) {
    hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(1), 0, stream, st, use_cfl, dt_override);
    return hipGetLastError();
}

hipError_t launch_step_end(StepState* st, hipStream_t stream) {
    hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(1), 0, stream, st);
    return hipGetLastError();
}

hipError_t launch_q_from_u(const HydroGeom& g, const double* U, double* Q, int fix, hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(q_from_u_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, U, Q, fix);
    return hipGetLastError();
}

hipError_t launch_u_from_q(const HydroGeom& g, double* Q, double* U, hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(u_from_q_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, Q, U);
    return hipGetLastError();
}

}  // namespace tg
