// hydro_device.hpp -- device helpers shared by the hydro stage kernels
// (FP64 parity kernel in hydro.hip, mixed-precision kernel in
// hydro_mixed.hip): tile geometry, floors, CFL term, error recording,
// block reduction and the ghost-cell mapping of the boundary conditions.
#pragma once

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

// fix_solution_mixed that also returns the primitives it rewrites U from
// (density and pressure after the floors, fractions, velocities): the same
// state QfromU would recover from the new U, without a second division and
// a second energy-minus-kinetic cancellation.
__device__ __forceinline__ bool fix_solution_mixed_q(const HydroGeom& g, double u[7], double q[7]) {
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
    q[DEN] = den;
    q[PRE] = pre;
    q[HII] = hii;
    q[ADV] = adv;
#pragma unroll
    for (int d = 0; d < 3; ++d) q[VEL + d] = v[d];
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
// dt of a stage or sub-step: explicit (>= 0) or from the step in flight
// (the DT_* codes).
__device__ __forceinline__ double stage_dtf(double dtf, const StepState* st, int stage) {
    if (dtf >= 0.0) return dtf;
    if (dtf == DT_HALF) return 0.5 * st->dt;
    if (dtf == DT_FULL) return st->dt;
    return stage == 1 ? st->dt / 2.0 : st->dt;
}

// Records error `bit` for the current step unless an earlier step failed
// first.  (Lane aggregation with a ballot and coherent pre-check loads here
// slowed the mixed stage kernel by half even with no error raised: a failing
// state instead ends the run at the next step, see step_begin_kernel.)
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

// Interior cell a stencil position at global (x, y, z) reads, which may lie
// in the 2-deep ghost layer of one dimension (the first out-of-range
// dimension decides), and the velocity rule of that ghost layer.
struct CellSrc {
    long long idx;
    int mode, mdim;  // mdim = 1 + ghost dimension, 0 for an interior cell
};

__device__ __forceinline__ CellSrc cell_source(const HydroGeom& g, int x, int y, int z) {
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
    return CellSrc{(long long)c[0] + (long long)g.nx * ((long long)c[1] + (long long)g.ny * c[2]), mode, mdim};
}

// Q of a stencil cell from its conserved values u: QfromU, in stage 1
// fixPrimitives and the "hydro before" check, then the ghost velocity rule.
template <int STAGE, bool MIXED>
__device__ __forceinline__ void q_of_cell(const HydroGeom& g, const double u[7], int mode, int mdim, double q[7],
                                          StepState* st, bool check) {
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

template <int STAGE, bool MIXED>
__device__ __forceinline__ void load_q(const HydroGeom& g, const double* __restrict__ U, int x, int y, int z,
                                       double q[7], StepState* st, bool check) {
    const CellSrc src = cell_source(g, x, y, z);
    double u[7];
#pragma unroll
    for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + src.idx];
    q_of_cell<STAGE, MIXED>(g, u, src.mode, src.mdim, q, st, check);
}

// Element (v, c) of a [7][ncell] state array.  A32: the variable's base
// (uniform, in scalar registers) plus a 32-bit byte offset, which the global
// loads and stores take as they are -- no 64-bit address arithmetic per
// access in vector registers; needs ncell * sizeof(T) < 2^32.
template <bool A32, typename T>
__device__ __forceinline__ T* soa_at(T* base, long long ncell, int v, long long c) {
    if constexpr (A32) return (T*)((char*)(base + v * ncell) + (unsigned int)((unsigned int)c * (unsigned int)sizeof(T)));
    else return base + v * ncell + c;
}


// The float primitives a stage reads, written per cell by the stage (or
// pass) that produced U: Q = QfromU(U) in FP64, rounded to float (not
// scaled).  In fix mode (the input of a predictor or single-stage sub-step)
// fixPrimitives is applied and Torch::checkValues("hydro before")
// (Torch.cpp:420-443) is evaluated on U: a failing cell is stored with a NaN
// density, which the consuming predictor reports at its step.
template <bool A32 = false>
__device__ __forceinline__ void store_qf(const HydroGeom& g, double q[7], const double u[7], float* __restrict__ qout,
                                         long long c, int fix, bool fixed_already) {
    if (fix) {
        if (!fixed_already) fix_primitives_mixed(g, q, u[DEN]);
        bool bad = (q[DEN] == 0.0) || (q[PRE] == 0.0);
#pragma unroll
        for (int v = 0; v < 7; ++v) bad |= (u[v] != u[v]) || isinf(u[v]);
        if (bad) q[DEN] = __builtin_nan("");
    }
#pragma unroll
    for (int v = 0; v < 7; ++v) *soa_at<A32>(qout, g.ncell, v, c) = (float)q[v];
}

// The float Q of a mixed-precision grid's cell for the next stage from its
// FP64 state (q_float_kernel in fix mode).
__device__ __forceinline__ void store_q_float(const HydroGeom& g, const double u[7], float* __restrict__ qout,
                                              long long c) {
    double q[7];
    q_from_u_mixed(q, u, g.gm1);
    store_qf(g, q, u, qout, c, 1, false);
}

}  // namespace tg
