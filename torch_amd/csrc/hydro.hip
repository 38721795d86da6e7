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
// This is the FP64 parity kernel: compiled with -ffp-contract=off it keeps
// the reference's operand order and is bitwise equal to its x86-64 build.
// The sweep fluxes of all three directions stay in LDS (140 KB) and are
// accumulated once per cell after the sweeps, so the accumulators are not
// live across the face solves.  The mixed-precision kernel is in
// hydro_mixed.hip.

#include "hydro_device.hpp"

namespace tg {

struct TileLDS {
    double qc[7][TCELLS];
    double qh[7][THALO];
    double fl[3][7][TFACES];  // face fluxes of the three sweeps
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

template <int DIM>
__device__ __forceinline__ void q_at(const TileLDS& s, int pos, int ta, int tb, double q[7]) {
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
template <int SOLVER, int DIM>
__device__ __forceinline__ bool face_flux(const HydroGeom& g, const TileLDS& s, int f, int ta, int tb,
                                          const GasConsts<double>& gc, double F[7]) {
    double ql[7], qr[7];
    if (g.spatial_order == 1) {
        double a[7], b[7], cc[7], dd[7];
        q_at<DIM>(s, f - 2, ta, tb, a);
        q_at<DIM>(s, f - 1, ta, tb, b);
        q_at<DIM>(s, f, ta, tb, cc);
        q_at<DIM>(s, f + 1, ta, tb, dd);
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            double dl = limiter<double>(g.limiter, b[v] - a[v], cc[v] - b[v]);
            ql[v] = b[v] + 0.5 * dl;   // QR of cell f-1
            double dr = limiter<double>(g.limiter, cc[v] - b[v], dd[v] - cc[v]);
            qr[v] = cc[v] - 0.5 * dr;  // QL of cell f
        }
    } else {
        q_at<DIM>(s, f - 1, ta, tb, ql);
        q_at<DIM>(s, f, ta, tb, qr);
    }
    double a_l2 = gc.g * ql[PRE] / ql[DEN];
    double a_r2 = gc.g * qr[PRE] / qr[DEN];
    return riemann<SOLVER, double, DIM>(F, ql, qr, a_l2, a_r2, gc);
}

template <int SOLVER, int STAGE, int DIM>
__device__ __forceinline__ void sweep(const HydroGeom& g, TileLDS& s, const double* __restrict__ Usrc,
                                      const int o[3], StepState* st, const GasConsts<double>& gc) {
    const int tid = threadIdx.x;
    const int N[3] = {g.nx, g.ny, g.nz};
    constexpr int DA = DIM == 0 ? 1 : 0;       // transverse dim "a"
    constexpr int DB = DIM == 2 ? 1 : 2;       // transverse dim "b"

    // 1. halo slabs of this direction (the previous sweep's face phase, the
    //    last reader of qh, ended at a barrier)
    if (tid < THALO) {
        const int h = tid >> 6, ta = tid & 7, tb = (tid >> 3) & 7;
        const int pos = h < 2 ? h - 2 : TILE + h - 2;
        int c[3];
        c[DIM] = o[DIM] + pos;
        c[DA] = o[DA] + ta;
        c[DB] = o[DB] + tb;
        double q[7];
        load_q<STAGE, false>(g, Usrc, c[0], c[1], c[2], q, st, false);
#pragma unroll
        for (int v = 0; v < 7; ++v) s.qh[v][tid] = q[v];
    }
    __syncthreads();

    // 2. one face per thread: face f sits between along-positions f-1 and f
    {
        const int f = tid % (TILE + 1), rest = tid / (TILE + 1);
        const int ta = rest & 7, tb = rest >> 3;
        const bool valid = (o[DIM] + f <= N[DIM]) && (o[DA] + ta < N[DA]) && (o[DB] + tb < N[DB]);
        double F[7];
        bool bad = false;
        if (valid) bad = face_flux<SOLVER, DIM>(g, s, f, ta, tb, gc, F);
        else {
#pragma unroll
            for (int v = 0; v < 7; ++v) F[v] = 0.0;
        }
        if (bad) set_error(st, STAGE == 1 ? ERR_RIEMANN_NAN_1 : ERR_RIEMANN_NAN_2);
#pragma unroll
        for (int v = 0; v < 7; ++v) s.fl[DIM][v][tid] = F[v];
    }
    __syncthreads();
}

// Flux divergence of one cell in the reference order, per dimension
// += (A/V) F_L, -= (A/V) F_R (Hydro.cpp:117-137).  The accumulators are only
// live here, not across the face solves.
__device__ __forceinline__ void flux_divergence(const HydroGeom& g, const TileLDS& s, int tid, double udot[7]) {
    const int lc[3] = {tid & 7, (tid >> 3) & 7, tid >> 6};
#pragma unroll
    for (int v = 0; v < 7; ++v) udot[v] = 0.0;
#pragma unroll
    for (int dim = 0; dim < 3; ++dim) {
        const int a = lc[dim];
        const int ta = lc[dim == 0 ? 1 : 0], tb = lc[dim == 2 ? 1 : 2];
        const int fL = (tb * TILE + ta) * (TILE + 1) + a;
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            double u = udot[v];
            u += g.kflux * s.fl[dim][v][fL];
            u -= g.kflux * s.fl[dim][v][fL + 1];
            udot[v] = u;
        }
    }
}

// dtf_host >= 0 overrides the stage's dt (st->dt / 2.0 for the predictor,
// st->dt for the corrector) with a host-given, already scaled dt.
template <int SOLVER, int STAGE>
__global__ __launch_bounds__(NTHREADS) void hydro_stage_kernel(HydroGeom g, const double* __restrict__ Usrc,
                                                              const double* Ubase, double* Udst, StepState* st,
                                                              double dtf_host) {
    __shared__ TileLDS s;
    if (st->halted) return;  // a halted sub-step's pass-through is launch_halted_copy's
    const int tile = tile_of_block(blockIdx.x, g.ntiles);
    const int tx = tile % g.ntx, ty = (tile / g.ntx) % g.nty, tz = tile / (g.ntx * g.nty);
    const int o[3] = {tx * TILE, ty * TILE, tz * TILE};
    const int tid = threadIdx.x;
    const GasConsts<double> gc = {g.g, g.gm1, g.g2};

    // tile cells: Q into LDS (ghost mapping for positions outside the grid)
    const int lx = tid & 7, ly = (tid >> 3) & 7, lz = tid >> 6;
    const int gx = o[0] + lx, gy = o[1] + ly, gz = o[2] + lz;
    const bool active = tid < TCELLS && gx < g.nx && gy < g.ny && gz < g.nz;
    if (tid < TCELLS) {
        double q[7];
        load_q<STAGE, false>(g, Usrc, gx, gy, gz, q, st, active);
#pragma unroll
        for (int v = 0; v < 7; ++v) s.qc[v][tid] = q[v];
    }
    sweep<SOLVER, STAGE, 0>(g, s, Usrc, o, st, gc);
    sweep<SOLVER, STAGE, 1>(g, s, Usrc, o, st, gc);
    sweep<SOLVER, STAGE, 2>(g, s, Usrc, o, st, gc);

    double inv_t = 0.0;
    if (active) {
        double udot[7];
        flux_divergence(g, s, tid, udot);
        // gravity source with GRAV = 0 (Hydro.cpp:134-137)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double grav = 0.0;
            udot[VEL + i] += grav;
            udot[PRE] += s.qc[VEL + i][tid] * grav;
        }
        const long long c = (long long)gx + (long long)g.nx * ((long long)gy + (long long)g.ny * gz);
        const double dtf = stage_dtf(dtf_host, st, STAGE);
        double u[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = Ubase[v * g.ncell + c];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] += dtf * udot[v];
        if (fix_solution(g, u)) set_error(st, STAGE == 1 ? ERR_FIX_SOLUTION_1 : ERR_FIX_SOLUTION_2);
#pragma unroll
        for (int v = 0; v < 7; ++v) Udst[v * g.ncell + c] = u[v];
        if (STAGE == 2) {
            double q[7];
            QfromU(q, u, g.gm1);
            fix_primitives(g, q, u[DEN]);
            inv_t = cfl_inv_t(g, q);
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

// Checkpointer::update's count for time t (Checkpointer.cpp:14-26): the
// number of the checkpoint percentages 100 / ncheckpoints passed.
__device__ __forceinline__ int checkpoint_count(const StepState* st, double t) {
    const double delta = 100.0 / st->ncheckpoints;
    const double pct = smin(100.0, 100.0 * t / st->tmax);
    int count = st->ckpt_count;
    while (pct - count * delta > delta - 1.0e-8) count += 1;
    return count;
}

__global__ void set_time_kernel(StepState* st, double t) {
    st->time = t;
    st->ckpt_count = 0;
    st->ckpt_count = checkpoint_count(st, t);
    st->paused = 0;
    st->halted = 0;
}

__global__ void resume_kernel(StepState* st) {
    if (st->paused) {
        st->paused = 0;
        st->halted = 0;
    }
}

// Start of a step: dt = min(dt_nextCheckpoint, dt_cfl) as Torch::run +
// Torch::fullStep + Torch::calculateTimeStep (Torch.cpp:266-270,304-339,397),
// with the checkpoint arithmetic of Checkpointer::update (Checkpointer.cpp:14-26).
// use_cfl = 0 takes dt_override instead of the CFL/first-step value.
__global__ void step_begin_kernel(StepState* st, int use_cfl, double dt_override,
                                  const unsigned long long* thermo_dt_bits,
                                  const unsigned long long* rad_dt_bits) {
    if (!use_cfl) {  // an explicit Torch::hydroStep(dt): no clock rules apply
        st->halted = 0;
        st->dt = dt_override;
        st->keep_ustar = 1;
        return;
    }
    if (st->halted) return;
    // a step that raised an error ends the run as the reference's exception
    // does; the error is reported by the next synchronising call; the loop
    // condition of Torch::run (Torch.cpp:266) also ends it once quitting
    if (!(st->time < st->tmax) || st->err || st->quitting) {
        st->halted = 1;
        return;
    }
    const int count = checkpoint_count(st, st->time);
    if (st->pause_ckpt && count > st->ckpt_count) {
        // a new checkpoint is due before this step: Torch::run prints here
        st->ckpt_count = count;
        st->paused = 1;
        st->halted = 1;
        return;
    }
    st->ckpt_count = count;
    const double delta = 100.0 / st->ncheckpoints;
    double pct = smin(100.0, 100.0 * st->time / st->tmax);
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
        // Torch::calculateTimeStep's debug test (Torch.cpp:321-332): its
        // condition reads the hydro percentage only; the step still runs and
        // Torch::run's loop ends after it
        if (st->debug && 100.0 * dt / st->tmax <= 1.0e-6) st->quitting = 1;
        if (thermo_dt_bits) {
            // Thermodynamics::calculateTimeStep: dt_max lowered by every
            // smaller |f E/RATE| (Thermodynamics.cpp:465-474)
            const double m = __longlong_as_double((long long)*thermo_dt_bits);
            dt = smin(dt, smin(m, st->dt_max));
        }
        if (rad_dt_bits) {
            // Radiation::calculateTimeStep's minimum, at most dt_max
            const double m = __longlong_as_double((long long)*rad_dt_bits);
            dt = smin(dt, smin(m, st->dt_max));
        }
    }
    st->dt = smin(dt_next, dt);
    st->maxinv = 0ull;
    // the next step_begin's tests on this step's end time (step_end adds dt
    // to time the same way): a pause for a new checkpoint, tmax, quitting
    const double t_end = st->time + st->dt;
    st->keep_ustar = checkpoint_count(st, t_end) > count || !(t_end < st->tmax) || st->quitting;
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
// cerr: the cooling operator's error word; a subcycle count beyond the int
// range raised by the integrate that produced `rate` is moved into the step
// state here (reported at this step, like the fixSolution errors) and the
// word cleared.
__global__ void apply_rate_fix_kernel(HydroGeom g, double* __restrict__ U, const double* __restrict__ rate, double dtf,
                                      StepState* st, unsigned int* cerr) {
    if (dtf < 0.0 && st->halted) return;
    if (cerr && blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned int e = *cerr;
        if (e & ERR_THERMO_NSTEPS) {
            set_error(st, ERR_THERMO_NSTEPS);
            *cerr = 0u;
        }
    }
    dtf = stage_dtf(dtf, st, 0);
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

// The cooling sub-step's update of U (as apply_rate_fix_kernel:
// Thermodynamics::updateSourceTerms + advSolution + fixSolution) fused with
// the float Q the next mixed stage reads (as q_float_kernel in fix mode), so
// the updated state is not read back.  A halted run leaves U as it is but
// still refreshes the float Q from it.  Compiled without FMA contraction,
// as the FP64 grids' update: the same arithmetic whichever kernel applies
// a cooling sub-step.
__global__ void apply_rate_fix_qf_kernel(HydroGeom g, double* __restrict__ U, const double* __restrict__ rate,
                                         double dtf, StepState* st, unsigned int* cerr, float* __restrict__ qout) {
    const bool halted = dtf < 0.0 && st->halted;
    if (!halted && cerr && blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned int e = *cerr;
        if (e & ERR_THERMO_NSTEPS) {
            set_error(st, ERR_THERMO_NSTEPS);
            *cerr = 0u;
        }
    }
    if (!halted) dtf = stage_dtf(dtf, st, 0);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long long)gridDim.x * blockDim.x) {
        double u[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + c];
        if (!halted) {
            double udot[7] = {0, 0, 0, 0, 0, 0, 0};
            udot[PRE] += rate ? rate[c] : 0.0;
#pragma unroll
            for (int v = 0; v < 7; ++v) u[v] += dtf * udot[v];
            if (fix_solution(g, u)) set_error(st, ERR_THERMO_FIX);
#pragma unroll
            for (int v = 0; v < 7; ++v) U[v * g.ncell + c] = u[v];
        }
        double q[7];
        q_from_u_mixed(q, u, g.gm1);
        store_qf(g, q, u, qout, c, 1, false);
    }
}

hipError_t launch_fix_solution_qf(const HydroGeom& g, double* U, const double* rate, double dtf, StepState* st,
                                  hipStream_t stream, unsigned int* cerr, float* qout) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(apply_rate_fix_qf_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, U, rate, dtf, st, cerr,
                       qout);
    return hipGetLastError();
}

// A halted run's single-stage sub-step (dt code DT_HALF / DT_FULL) still
// passes its input through to its output, float primitives included: the
// host swaps the two state buffers after every such sub-step without knowing
// whether the run halted.  Launched after each such sub-step; returns at
// once while the run is going.
__global__ void halted_copy_kernel(HydroGeom g, const StepState* st, const double* __restrict__ Usrc,
                                   double* __restrict__ Udst, const float* __restrict__ qin,
                                   float* __restrict__ qout) {
    if (!st->halted) return;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int v = 0; v < 7; ++v) Udst[v * g.ncell + c] = Usrc[v * g.ncell + c];
        if (qin) {
#pragma unroll
            for (int v = 0; v < 7; ++v) qout[v * g.ncell + c] = qin[v * g.ncell + c];
        }
    }
}

__global__ void zero_kernel(double* __restrict__ p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

hipError_t launch_zero(double* p, long long n, hipStream_t stream) {
    long long nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)nb), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launch_halted_copy(const HydroGeom& g, const StepState* st, const double* Usrc, double* Udst,
                              const float* qin, float* qout, hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(halted_copy_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, st, Usrc, Udst, qin, qout);
    return hipGetLastError();
}

hipError_t launch_fix_solution(const HydroGeom& g, double* U, const double* rate, double dtf, StepState* st,
                               hipStream_t stream, unsigned int* cerr) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(apply_rate_fix_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, U, rate, dtf, st, cerr);
    return hipGetLastError();
}

template <int SOLVER>
static hipError_t launch_stage_t(const HydroGeom& g, int stage, const double* Usrc, const double* Ubase, double* Udst,
                                 StepState* st, double dtf, hipStream_t stream) {
    dim3 grid(g.ntiles), block(NTHREADS);
    if (stage == 1)
        hipLaunchKernelGGL((hydro_stage_kernel<SOLVER, 1>), grid, block, 0, stream, g, Usrc, Ubase, Udst, st, dtf);
    else
        hipLaunchKernelGGL((hydro_stage_kernel<SOLVER, 2>), grid, block, 0, stream, g, Usrc, Ubase, Udst, st, dtf);
    return hipGetLastError();
}

hipError_t launch_hydro_stage(const HydroGeom& g, int solver, int stage, const double* Usrc, const double* Ubase,
                              double* Udst, StepState* st, hipStream_t stream, double dtf) {
    switch (solver) {
    case SOLVER_HLL:  return launch_stage_t<SOLVER_HLL>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    case SOLVER_HLLC: return launch_stage_t<SOLVER_HLLC>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    default:          return launch_stage_t<SOLVER_ROTATED>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    }
}

hipError_t launch_cfl(const HydroGeom& g, const double* U, StepState* st, hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(cfl_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, U, st);
    return hipGetLastError();
}

hipError_t launch_step_begin(StepState* st, int use_cfl, double dt_override, hipStream_t stream,
                             const unsigned long long* thermo_dt_bits, const unsigned long long* rad_dt_bits) {
    hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(1), 0, stream, st, use_cfl, dt_override, thermo_dt_bits,
                       rad_dt_bits);
    return hipGetLastError();
}

// dst = src (n doubles) unless the run has halted.
__global__ void copy_if_running_kernel(const StepState* st, const double* __restrict__ src, double* __restrict__ dst,
                                       long long n) {
    if (st->halted) return;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy_if_running(const StepState* st, const double* src, double* dst, long long n,
                                  hipStream_t stream) {
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(copy_if_running_kernel, dim3((unsigned)nb), dim3(256), 0, stream, st, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_set_time(StepState* st, double t, hipStream_t stream) {
    hipLaunchKernelGGL(set_time_kernel, dim3(1), dim3(1), 0, stream, st, t);
    return hipGetLastError();
}

hipError_t launch_resume(StepState* st, hipStream_t stream) {
    hipLaunchKernelGGL(resume_kernel, dim3(1), dim3(1), 0, stream, st);
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

__global__ void q_from_u_if_running_kernel(HydroGeom g, const StepState* st, const double* __restrict__ U,
                                           double* __restrict__ Q) {
    if (st->halted) return;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long long)gridDim.x * blockDim.x) {
        double u[7], q[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * g.ncell + c];
        QfromU(q, u, g.gm1);
        fix_primitives(g, q, u[DEN]);
#pragma unroll
        for (int v = 0; v < 7; ++v) Q[v * g.ncell + c] = q[v];
    }
}

hipError_t launch_q_from_u_if_running(const HydroGeom& g, const StepState* st, const double* U, double* Q,
                                      hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(q_from_u_if_running_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, st, U, Q);
    return hipGetLastError();
}

hipError_t launch_u_from_q(const HydroGeom& g, double* Q, double* U, hipStream_t stream) {
    long long nb = (g.ncell + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(u_from_q_kernel, dim3((unsigned)nb), dim3(256), 0, stream, g, Q, U);
    return hipGetLastError();
}

}  // namespace tg
