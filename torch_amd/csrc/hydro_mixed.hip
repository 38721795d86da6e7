// hydro_mixed.hip -- mixed-precision hydro stage kernel for CDNA4 (gfx950).
//
// The same stage as hydro.hip's FP64 kernel (Torch::hydroStep stage, see
// there for the reference map), with reconstruction and Riemann fluxes in
// float and the state, the flux accumulation, the primitive recovery, the
// floors and the CFL reduction in FP64.
//
// Layout for throughput: a workgroup (576 threads, 9 waves) owns an 8x8x8
// tile.  One LDS array q[7][1280] (float, one plane per variable) holds the
// tile's 512 cells and the 2-deep halo slabs of all three directions
// (3 x 4 x 64 cells), loaded once at the start by all threads together; a
// face finds each stencil cell with a select between its tile and halo
// index (no branches) and reads the 7 variables with one address and
// immediate plane offsets.  Per direction every thread solves one face into
// an LDS flux array and the 512 cell threads accumulate their two faces
// into FP64 registers in the reference's order.  For the y and z sweeps
// consecutive lanes take consecutive x; for the x sweep consecutive lanes
// walk along x.  LDS 52 KB, 74 VGPRs (six waves per SIMD: two workgroups
// per CU; capping the registers at the 72 that three workgroups need
// spills and measured 6 % slower).
//
// FP32 range: the float state is the FP64 state scaled by powers of two,
// density by 2^kr and pressure by 2^kp with the grid's largest density and
// pressure brought near 2^16, velocities by 2^kv with 2 kv = kp - kr (so
// that the Euler relations hold in the scaled units), HII and ADV fractions
// unscaled.  Slope squares, S_l S_r dU and the energy fluxes then stay in
// FP32 range whatever the code units: with the reference's own scalings
// (1e-20 g/cm^3, 1e-28 dyn/cm^2, 1e11 s) pressures are ~1e22 and their
// squares would overflow.  Scaling by powers of two is exact, the fluxes are
// unscaled exactly in the FP64 accumulation (A/V 2^-k per variable), and
// the reference's thresholds (limiter a b <= 1e-30, |velocity jump| < 1e-6)
// are scaled alike.  Every predictor launch records its input's largest
// density and pressure (float bit patterns, atomicMax) for the next launch's
// scale; set_primitive / set_conserved / reset_clock seed it with a
// reduction over the grid.
#include "hydro_device.hpp"

namespace tg {

constexpr int NLOAD = TCELLS + 3 * THALO;        // tile cells + halo slabs: 1280

struct MixedLDS {
    float q[7][NLOAD];   // [0, 512): tile cells x + 8 (y + 8 z); [512, 1280): halo slabs
    float fl[7][TFACES];
    double red[NTHREADS / 64];
    unsigned int qmw[NTHREADS / 64][2];  // per-wave largest density / pressure (float bits)
};

// Power-of-two state scale of a launch (see the header) and the reference's
// thresholds in the scaled units.  Everything here is uniform over the
// launch and derived with integer arithmetic, so it stays in scalar registers.
struct MixedScale {
    int kr, kp, kv;
    float thr_r, thr_p, thr_v;  // limiter threshold 1e-30 for density, pressure, velocity
    float nn_thr;               // |velocity jump|^2 threshold 1e-12
};

// 2^k brings the largest value (float bits) into [2^15, 2^16); no value
// recorded (0) or a non-finite one leaves the variable unscaled.
__device__ __forceinline__ int scale_shift(unsigned int bits) {
    const int e = (int)(bits >> 23) - 127;
    if (bits == 0 || e > 127) return 0;
    const int k = 15 - e;
    return k < -100 ? -100 : (k > 100 ? 100 : k);
}

// x 2^k2 for a positive normal float x given by its bits, the exponent
// clamped to the normal range (integer min/max: scalar instructions).
__device__ __forceinline__ float scaled_threshold(unsigned int bits, int k2) {
    const int e = min(max((int)(bits >> 23) + k2, 1), 254);
    return __uint_as_float((bits & 0x007fffffu) | ((unsigned int)e << 23));
}

__device__ __forceinline__ MixedScale mixed_scale(const StepState* st) {
    MixedScale m;
    m.kr = scale_shift(st->qmax[0]);
    m.kp = scale_shift(st->qmax[1]);
    if ((m.kp - m.kr) & 1) m.kp -= 1;
    m.kv = (m.kp - m.kr) / 2;
    const unsigned int t30 = __float_as_uint(1.0e-30f), t12 = __float_as_uint(1.0e-12f);
    m.thr_r = scaled_threshold(t30, 2 * m.kr);
    m.thr_p = scaled_threshold(t30, 2 * m.kp);
    m.thr_v = scaled_threshold(t30, 2 * m.kv);
    m.nn_thr = scaled_threshold(t12, 2 * m.kv);
    return m;
}

__device__ __forceinline__ float thr_of(const MixedScale& m, int v) {
    return v == DEN ? m.thr_r : (v == PRE ? m.thr_p : (v >= VEL ? m.thr_v : 1.0e-30f));
}

__device__ __forceinline__ unsigned int max_bits(double v) {
    return v > 0.0 ? __float_as_uint((float)fmin(v, 3.0e38)) : 0u;
}

// LDS slot of along-position pos (-2..9) of pencil (ta, tb) in sweep DIM.
// Tile cells: x + 8 (y + 8 z); halo cell of direction DIM at layer h
// (pos -2, -1, 8, 9 -> h 0..3): 512 + 256 DIM + 64 h + 8 tb + ta.
template <int DIM>
__device__ __forceinline__ int slot(int pos, int ta, int tb) {
    int c;
    if (DIM == 0) c = pos + TILE * (ta + TILE * tb);
    else if (DIM == 1) c = ta + TILE * (pos + TILE * tb);
    else c = ta + TILE * (tb + TILE * pos);
    const int h = pos < 0 ? pos + 2 : pos - (TILE - 2);
    const int hs = TCELLS + DIM * THALO + h * (TILE * TILE) + tb * TILE + ta;
    return (unsigned)pos < (unsigned)TILE ? c : hs;
}

// Face index of along-position f of pencil (ta, tb) in sweep DIM.
template <int DIM>
__device__ __forceinline__ int face_id(int f, int ta, int tb) {
    if (DIM == 0) return (tb * TILE + ta) * (TILE + 1) + f;
    return f * (TILE * TILE) + tb * TILE + ta;
}

template <int DIM>
__device__ __forceinline__ void read_q(const MixedLDS& s, int pos, int ta, int tb, float q[7]) {
    const int i = slot<DIM>(pos, ta, tb);
#pragma unroll
    for (int v = 0; v < 7; ++v) q[v] = s.q[v][i];
}

template <int SOLVER, bool ALB, int DIM>
__device__ __forceinline__ bool face_flux_mixed(const HydroGeom& g, const MixedLDS& s, const MixedScale& m, int f,
                                                int ta, int tb, const GasConsts<float>& gc, float F[7]) {
    float ql[7], qr[7];
    if (g.spatial_order == 1) {
        float a[7], b[7], cc[7], dd[7];
        read_q<DIM>(s, f - 2, ta, tb, a);
        read_q<DIM>(s, f - 1, ta, tb, b);
        read_q<DIM>(s, f, ta, tb, cc);
        read_q<DIM>(s, f + 1, ta, tb, dd);
        if constexpr (ALB) {
#pragma unroll
            for (int v = 0; v < 7; ++v) {
                ql[v] = b[v] + 0.5f * albada_f(b[v] - a[v], cc[v] - b[v], thr_of(m, v));
                qr[v] = cc[v] - 0.5f * albada_f(cc[v] - b[v], dd[v] - cc[v], thr_of(m, v));
            }
        } else {
#pragma unroll
            for (int v = 0; v < 7; ++v) {
                ql[v] = b[v] + 0.5f * limiter_f(g.limiter, b[v] - a[v], cc[v] - b[v], thr_of(m, v));
                qr[v] = cc[v] - 0.5f * limiter_f(g.limiter, cc[v] - b[v], dd[v] - cc[v], thr_of(m, v));
            }
        }
    } else {
        read_q<DIM>(s, f - 1, ta, tb, ql);
        read_q<DIM>(s, f, ta, tb, qr);
    }
    return riemann_f<SOLVER, DIM>(F, ql, qr, gc, g.igm1f, m.nn_thr);
}

template <int SOLVER, bool ALB, int STAGE, int DIM>
__device__ __forceinline__ void sweep_mixed(const HydroGeom& g, MixedLDS& s, const MixedScale& m, const int o[3],
                                            StepState* st, const GasConsts<float>& gc, double udot[7]) {
    const int tid = threadIdx.x;
    const int N[3] = {g.nx, g.ny, g.nz};
    constexpr int DA = DIM == 0 ? 1 : 0;
    constexpr int DB = DIM == 2 ? 1 : 2;
    {
        int f, ta, tb;
        if (DIM == 0) {
            f = tid % (TILE + 1);
            const int rest = tid / (TILE + 1);
            ta = rest & 7;
            tb = rest >> 3;
        } else {
            f = tid >> 6;
            ta = tid & 7;
            tb = (tid >> 3) & 7;
        }
        const bool valid = (o[DIM] + f <= N[DIM]) && (o[DA] + ta < N[DA]) && (o[DB] + tb < N[DB]);
        float F[7];
        bool bad = false;
        if (valid) {
            bad = face_flux_mixed<SOLVER, ALB, DIM>(g, s, m, f, ta, tb, gc, F);
        } else {
#pragma unroll
            for (int v = 0; v < 7; ++v) F[v] = 0.f;
        }
        if (bad) set_error(st, STAGE == 1 ? ERR_RIEMANN_NAN_1 : ERR_RIEMANN_NAN_2);
        const int id = face_id<DIM>(f, ta, tb);
#pragma unroll
        for (int v = 0; v < 7; ++v) s.fl[v][id] = F[v];
    }
    __syncthreads();
    // += (A/V) F_L, -= (A/V) F_R (Hydro.cpp:117-137)
    if (tid < TCELLS) {
        const int lc[3] = {tid & 7, (tid >> 3) & 7, tid >> 6};
        const int fL = face_id<DIM>(lc[DIM], lc[DA], lc[DB]);
        const int fR = face_id<DIM>(lc[DIM] + 1, lc[DA], lc[DB]);
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            double u = udot[v];
            u += g.kflux * (double)s.fl[v][fL];
            u -= g.kflux * (double)s.fl[v][fR];
            udot[v] = u;
        }
    }
    __syncthreads();
}

// Tile-relative coordinates of load slot idx: the tile cell x + 8 (y + 8 z),
// or halo slot 512 + 256 dim + 64 h + 8 tb + ta.
__device__ __forceinline__ void slot_coords(int idx, int c[3]) {
    if (idx < TCELLS) {
        c[0] = idx & 7;
        c[1] = (idx >> 3) & 7;
        c[2] = idx >> 6;
    } else {
        const int k = idx - TCELLS;
        const int dim = k >> 8, r = k & 255;
        const int h = r >> 6, ta = r & 7, tb = (r >> 3) & 7;
        const int pos = h < 2 ? h - 2 : TILE + h - 2;
        if (dim == 0) { c[0] = pos; c[1] = ta; c[2] = tb; }
        else if (dim == 1) { c[0] = ta; c[1] = pos; c[2] = tb; }
        else { c[0] = ta; c[1] = tb; c[2] = pos; }
    }
}

// Ghost-layer velocity rule (see flip_normal) on float primitives.
__device__ __forceinline__ void flip_normal_f(float& vn, int mode) {
    switch (mode) {
    case 1: vn = -vn; break;
    case 2: if (vn > 0.f) vn = -vn; break;
    case 3: if (vn < 0.f) vn = -vn; break;
    case 4: if (vn < 0.f) vn = -vn; break;
    case 5: if (vn > 0.f) vn = -vn; break;
    default: break;
    }
}

__device__ __forceinline__ void apply_ghost_mode_f(float q[7], int mode, int dim) {
    if (dim == 0) flip_normal_f(q[VEL + 0], mode);
    else if (dim == 1) flip_normal_f(q[VEL + 1], mode);
    else flip_normal_f(q[VEL + 2], mode);
}

__device__ __forceinline__ unsigned int max_bits_f(float v) { return v > 0.f ? __float_as_uint(v) : 0u; }

// The float primitives a stage reads, written per cell by the stage (or
// pass) that produced U: Q = QfromU(U) in FP64, rounded to float (not
// scaled).  In fix mode (the input of a predictor or single-stage sub-step)
// fixPrimitives is applied and Torch::checkValues("hydro before")
// (Torch.cpp:420-443) is evaluated on U: a failing cell is stored with a NaN
// density, which the consuming predictor reports at its step.
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
    for (int v = 0; v < 7; ++v) qout[v * g.ncell + c] = (float)q[v];
}

// The tile cells and the halo slabs of all three directions, three rounds of
// 576 slots: float Q of the source cell, ghost velocity rule, power-of-two
// scale, into LDS.  INNER: the tile and its 2-deep halos lie inside the grid
// (no ghost cells, no clamping).  (Issuing all three rounds' loads before
// using any measured slower: the extra registers cost more than the overlap
// gains.)
template <int STAGE, bool INNER>
__device__ __forceinline__ void load_tile(const HydroGeom& g, const float* __restrict__ qin, const int o[3],
                                          MixedLDS& s, const int ks[7], StepState* st, unsigned int& mr,
                                          unsigned int& mp) {
    constexpr int NR = (NLOAD + NTHREADS - 1) / NTHREADS;
    const int tid = threadIdx.x;
    for (int r = 0; r < NR; ++r) {
        const int idx = tid + r * NTHREADS;
        if (idx < NLOAD) {
            int c[3];
            slot_coords(idx, c);
            long long src;
            int mode = 0, mdim = 0;
            if (INNER) {
                src = (long long)(o[0] + c[0]) +
                      (long long)g.nx * ((long long)(o[1] + c[1]) + (long long)g.ny * (o[2] + c[2]));
            } else {
                const CellSrc cs = cell_source(g, o[0] + c[0], o[1] + c[1], o[2] + c[2]);
                src = cs.idx;
                mode = cs.mode;
                mdim = cs.mdim;
            }
            float q[7];
#pragma unroll
            for (int v = 0; v < 7; ++v) q[v] = qin[v * g.ncell + src];
            if (mdim) apply_ghost_mode_f(q, mode, mdim - 1);
            if (STAGE == 1 && r == 0 && idx < TCELLS) {
                const bool check = INNER || ((o[0] + (idx & 7) < g.nx) && (o[1] + ((idx >> 3) & 7) < g.ny) &&
                                             (o[2] + (idx >> 6) < g.nz));
                if (check) {
                    if (q[DEN] != q[DEN]) set_error(st, ERR_CHECK_VALUES);
                    mr = max_bits_f(q[DEN]);
                    mp = max_bits_f(q[PRE]);
                }
            }
#pragma unroll
            for (int v = 0; v < 7; ++v) s.q[v][idx] = __builtin_ldexpf(q[v], ks[v]);
        }
    }
}

// Six waves per SIMD (at most 80 VGPRs): two 9-wave workgroups per CU.
// ALB: van Albada limiter (the other limiters in the ALB = false instance).
// qin: float Q of the input state (fixPrimitives applied for stage 1);
// qout: float Q of the output state for the next stage, fixed if qout_fix.
template <int SOLVER, bool ALB, int STAGE>
__global__ __launch_bounds__(NTHREADS) __attribute__((amdgpu_waves_per_eu(7))) void hydro_stage_mixed_kernel(
    HydroGeom g, const float* __restrict__ qin, const double* Ubase, double* Udst, float* __restrict__ qout,
    int qout_fix, StepState* st, double dtf_host) {
    __shared__ MixedLDS s;
    // A halted sub-step's pass-through is launch_halted_copy's: a copy branch
    // here (never taken while running) made every stage 45% slower at 256^3.
    if (st->halted) return;
    const int tile = tile_of_block(blockIdx.x, g.ntiles);
    const int tx = tile % g.ntx, ty = (tile / g.ntx) % g.nty, tz = tile / (g.ntx * g.nty);
    const int o[3] = {tx * TILE, ty * TILE, tz * TILE};
    const int tid = threadIdx.x;
    const GasConsts<float> gc = {g.gf, g.gm1f, g.g2f};
    const MixedScale m = mixed_scale(st);
    const int ks[7] = {m.kr, m.kp, 0, 0, m.kv, m.kv, m.kv};

    const int lx = tid & 7, ly = (tid >> 3) & 7, lz = tid >> 6;
    const int gx = o[0] + lx, gy = o[1] + ly, gz = o[2] + lz;
    const bool active = tid < TCELLS && gx < g.nx && gy < g.ny && gz < g.nz;
    const long long cell = (long long)gx + (long long)g.nx * ((long long)gy + (long long)g.ny * gz);

    unsigned int mr = 0, mp = 0;
    const bool inner = o[0] >= 2 && o[1] >= 2 && o[2] >= 2 && o[0] + TILE + 2 <= g.nx &&
                       o[1] + TILE + 2 <= g.ny && o[2] + TILE + 2 <= g.nz;
    if (inner) load_tile<STAGE, true>(g, qin, o, s, ks, st, mr, mp);
    else load_tile<STAGE, false>(g, qin, o, s, ks, st, mr, mp);
    if (STAGE == 1) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mr = max(mr, (unsigned int)__shfl_xor((int)mr, off, 64));
            mp = max(mp, (unsigned int)__shfl_xor((int)mp, off, 64));
        }
        if ((tid & 63) == 0) {
            s.qmw[tid >> 6][0] = mr;
            s.qmw[tid >> 6][1] = mp;
        }
    }
    __syncthreads();
    if (STAGE == 1 && tid == 0) {
        unsigned int a = 0, b = 0;
        for (int w = 0; w < NTHREADS / 64; ++w) {
            a = max(a, s.qmw[w][0]);
            b = max(b, s.qmw[w][1]);
        }
        if (a) atomicMax(&st->qmax_next[0], a);
        if (b) atomicMax(&st->qmax_next[1], b);
    }

    double udot[7] = {0, 0, 0, 0, 0, 0, 0};
    sweep_mixed<SOLVER, ALB, STAGE, 0>(g, s, m, o, st, gc, udot);
    sweep_mixed<SOLVER, ALB, STAGE, 1>(g, s, m, o, st, gc, udot);
    sweep_mixed<SOLVER, ALB, STAGE, 2>(g, s, m, o, st, gc, udot);
    {
        // the sums are in flux units scaled by 2^(kr+kv) (mass, HII, ADV),
        // 2^(kp+kv) (energy) and 2^kp (momentum): unscaled exactly
        const int ke[7] = {m.kr + m.kv, m.kp + m.kv, m.kr + m.kv, m.kr + m.kv, m.kp, m.kp, m.kp};
#pragma unroll
        for (int v = 0; v < 7; ++v) udot[v] = ldexp(udot[v], -ke[v]);
    }

    double inv_t = 0.0;
    if (active) {
        // gravity source with GRAV = 0 (Hydro.cpp:134-137) adds nothing
        const long long c = cell;
        const double dtf = stage_dtf(dtf_host, st, STAGE);
        double u[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = Ubase[v * g.ncell + c];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] += dtf * udot[v];
        if (fix_solution_mixed(g, u)) set_error(st, STAGE == 1 ? ERR_FIX_SOLUTION_1 : ERR_FIX_SOLUTION_2);
#pragma unroll
        for (int v = 0; v < 7; ++v) Udst[v * g.ncell + c] = u[v];
        double q[7];
        q_from_u_mixed(q, u, g.gm1);
        if (STAGE == 2) {
            fix_primitives_mixed(g, q, u[DEN]);
            inv_t = cfl_inv_t_mixed(g, q);
        }
        store_qf(g, q, u, qout, c, qout_fix, STAGE == 2);
    }
    if (STAGE == 2) block_max_to(inv_t, &st->maxinv, s.red);
}

template <int SOLVER, bool ALB>
static hipError_t launch_mixed_t(const HydroGeom& g, int stage, const float* qin, const double* Ubase, double* Udst,
                                 float* qout, int qout_fix, StepState* st, double dtf, hipStream_t stream) {
    dim3 grid(g.ntiles), block(NTHREADS);
    if (stage == 1)
        hipLaunchKernelGGL((hydro_stage_mixed_kernel<SOLVER, ALB, 1>), grid, block, 0, stream, g, qin, Ubase, Udst,
                           qout, qout_fix, st, dtf);
    else
        hipLaunchKernelGGL((hydro_stage_mixed_kernel<SOLVER, ALB, 2>), grid, block, 0, stream, g, qin, Ubase, Udst,
                           qout, qout_fix, st, dtf);
    return hipGetLastError();
}

// The van Albada instance carries no code of the other limiters: the
// generic limiter switch at the 42 reconstruction sites (14 per sweep)
// triples the kernel's code (57 KB against 19 KB), close to the 64 KB
// instruction cache two CUs share.
template <int SOLVER>
static hipError_t launch_mixed_l(const HydroGeom& g, int stage, const float* qin, const double* Ubase, double* Udst,
                                 float* qout, int qout_fix, StepState* st, double dtf, hipStream_t stream) {
    if (g.limiter == LIM_ALBADA)
        return launch_mixed_t<SOLVER, true>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, stream);
    return launch_mixed_t<SOLVER, false>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, stream);
}

// Float Q of a state for the next mixed stage (see store_qf), for a state
// that no stage produced (set by the caller, or changed by another operator).
__global__ __launch_bounds__(256) void q_float_kernel(HydroGeom g, const double* __restrict__ U, float* __restrict__ q,
                                                      int fix) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell;
         c += (long long)gridDim.x * blockDim.x) {
        double u[7], qd[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + c];
        q_from_u_mixed(qd, u, g.gm1);
        store_qf(g, qd, u, q, c, fix, false);
    }
}

hipError_t launch_q_float(const HydroGeom& g, const double* U, float* q, int fix, hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(q_float_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, U, q, fix);
    return hipGetLastError();
}

// Before a predictor launch: the largest density and pressure the previous
// predictor (or the seeding reduction) recorded become the scale; the record
// starts afresh.  No record keeps the scale in use.
__global__ void mixed_scale_rotate_kernel(StepState* st) {
    if (st->qmax_next[0] | st->qmax_next[1]) {
        st->qmax[0] = st->qmax_next[0];
        st->qmax[1] = st->qmax_next[1];
    }
    st->qmax_next[0] = 0;
    st->qmax_next[1] = 0;
}

// Largest density and pressure (after the floors of fixPrimitives) over the
// grid into qmax_next, one atomicMax per wave.
__global__ __launch_bounds__(256) void state_max_kernel(HydroGeom g, const double* __restrict__ U, StepState* st) {
    unsigned int mr = 0, mp = 0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell;
         c += (long long)gridDim.x * blockDim.x) {
        double u[7], q[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + c];
        q_from_u_mixed(q, u, g.gm1);
        fix_primitives_mixed(g, q, u[DEN]);
        mr = max(mr, max_bits(q[DEN]));
        mp = max(mp, max_bits(q[PRE]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mr = max(mr, (unsigned int)__shfl_xor((int)mr, off, 64));
        mp = max(mp, (unsigned int)__shfl_xor((int)mp, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        if (mr) atomicMax(&st->qmax_next[0], mr);
        if (mp) atomicMax(&st->qmax_next[1], mp);
    }
}

hipError_t launch_mixed_scale_seed(const HydroGeom& g, const double* U, StepState* st, hipStream_t stream) {
    // clear the record, reduce the grid into it, make it the scale
    hipLaunchKernelGGL(mixed_scale_rotate_kernel, dim3(1), dim3(1), 0, stream, st);
    long long nb = (g.ncell + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(state_max_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, U, st);
    hipLaunchKernelGGL(mixed_scale_rotate_kernel, dim3(1), dim3(1), 0, stream, st);
    return hipGetLastError();
}

hipError_t launch_hydro_stage_mixed(const HydroGeom& g, int solver, int stage, const float* qin, const double* Ubase,
                                    double* Udst, float* qout, int qout_fix, StepState* st, hipStream_t stream,
                                    double dtf) {
    if (stage == 1) hipLaunchKernelGGL(mixed_scale_rotate_kernel, dim3(1), dim3(1), 0, stream, st);
    switch (solver) {
    case SOLVER_HLL:  return launch_mixed_l<SOLVER_HLL>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, stream);
    case SOLVER_HLLC: return launch_mixed_l<SOLVER_HLLC>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, stream);
    default:
        return launch_mixed_l<SOLVER_ROTATED>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, stream);
    }
}

}  // namespace tg
