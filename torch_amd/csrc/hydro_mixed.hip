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
// walk along x.  LDS 52 KB: three workgroups per CU.
#include "hydro_device.hpp"

namespace tg {

constexpr int NLOAD = TCELLS + 3 * THALO;        // tile cells + halo slabs: 1280

struct MixedLDS {
    float q[7][NLOAD];   // [0, 512): tile cells x + 8 (y + 8 z); [512, 1280): halo slabs
    float fl[7][TFACES];
    double red[NTHREADS / 64];
};

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

template <int SOLVER, int DIM>
__device__ __forceinline__ bool face_flux_mixed(const HydroGeom& g, const MixedLDS& s, int f, int ta, int tb,
                                                const GasConsts<float>& gc, float F[7]) {
    float ql[7], qr[7];
    if (g.spatial_order == 1) {
        float a[7], b[7], cc[7], dd[7];
        read_q<DIM>(s, f - 2, ta, tb, a);
        read_q<DIM>(s, f - 1, ta, tb, b);
        read_q<DIM>(s, f, ta, tb, cc);
        read_q<DIM>(s, f + 1, ta, tb, dd);
        if (g.limiter == LIM_ALBADA) {
#pragma unroll
            for (int v = 0; v < 7; ++v) {
                ql[v] = b[v] + 0.5f * albada_f(b[v] - a[v], cc[v] - b[v]);
                qr[v] = cc[v] - 0.5f * albada_f(cc[v] - b[v], dd[v] - cc[v]);
            }
        } else {
#pragma unroll
            for (int v = 0; v < 7; ++v) {
                ql[v] = b[v] + 0.5f * limiter<float>(g.limiter, b[v] - a[v], cc[v] - b[v]);
                qr[v] = cc[v] - 0.5f * limiter<float>(g.limiter, cc[v] - b[v], dd[v] - cc[v]);
            }
        }
    } else {
        read_q<DIM>(s, f - 1, ta, tb, ql);
        read_q<DIM>(s, f, ta, tb, qr);
    }
    return riemann_f<SOLVER, DIM>(F, ql, qr, gc, g.igm1f);
}

template <int SOLVER, int STAGE, int DIM>
__device__ __forceinline__ void sweep_mixed(const HydroGeom& g, MixedLDS& s, const int o[3], StepState* st,
                                            const GasConsts<float>& gc, double udot[7]) {
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
            bad = face_flux_mixed<SOLVER, DIM>(g, s, f, ta, tb, gc, F);
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

template <int SOLVER, int STAGE>
__global__ __launch_bounds__(NTHREADS) void hydro_stage_mixed_kernel(HydroGeom g, const double* __restrict__ Usrc,
                                                                    const double* Ubase, double* Udst, StepState* st,
                                                                    double dtf_host) {
    __shared__ MixedLDS s;
    if (st->halted) return;
    const int tile = tile_of_block(blockIdx.x, g.ntiles);
    const int tx = tile % g.ntx, ty = (tile / g.ntx) % g.nty, tz = tile / (g.ntx * g.nty);
    const int o[3] = {tx * TILE, ty * TILE, tz * TILE};
    const int tid = threadIdx.x;
    const GasConsts<float> gc = {g.gf, g.gm1f, g.g2f};

    // tile cells and the halo slabs of all three directions: Q (FP64
    // recovery, floors in stage 1) rounded to float into LDS
    for (int idx = tid; idx < NLOAD; idx += NTHREADS) {
        int c[3];
        bool check = false;
        if (idx < TCELLS) {
            c[0] = idx & 7;
            c[1] = (idx >> 3) & 7;
            c[2] = idx >> 6;
            check = (o[0] + c[0] < g.nx) && (o[1] + c[1] < g.ny) && (o[2] + c[2] < g.nz);
        } else {
            const int k = idx - TCELLS;
            const int dim = k >> 8, r = k & 255;
            const int h = r >> 6, ta = r & 7, tb = (r >> 3) & 7;
            const int pos = h < 2 ? h - 2 : TILE + h - 2;
            if (dim == 0) { c[0] = pos; c[1] = ta; c[2] = tb; }
            else if (dim == 1) { c[0] = ta; c[1] = pos; c[2] = tb; }
            else { c[0] = ta; c[1] = tb; c[2] = pos; }
        }
        double q[7];
        load_q<STAGE, true>(g, Usrc, o[0] + c[0], o[1] + c[1], o[2] + c[2], q, st, check);
        // slot idx: the tile cell x + 8 (y + 8 z), or halo slot 512 + 256 dim + 64 h + 8 tb + ta
#pragma unroll
        for (int v = 0; v < 7; ++v) s.q[v][idx] = (float)q[v];
    }
    __syncthreads();

    double udot[7] = {0, 0, 0, 0, 0, 0, 0};
    sweep_mixed<SOLVER, STAGE, 0>(g, s, o, st, gc, udot);
    sweep_mixed<SOLVER, STAGE, 1>(g, s, o, st, gc, udot);
    sweep_mixed<SOLVER, STAGE, 2>(g, s, o, st, gc, udot);

    const int lx = tid & 7, ly = (tid >> 3) & 7, lz = tid >> 6;
    const int gx = o[0] + lx, gy = o[1] + ly, gz = o[2] + lz;
    const bool active = tid < TCELLS && gx < g.nx && gy < g.ny && gz < g.nz;
    double inv_t = 0.0;
    if (active) {
        // gravity source with GRAV = 0 (Hydro.cpp:134-137) adds nothing
        const long long c = (long long)gx + (long long)g.nx * ((long long)gy + (long long)g.ny * gz);
        const double dtf = dtf_host >= 0.0 ? dtf_host : (STAGE == 1 ? st->dt / 2.0 : st->dt);
        double u[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = Ubase[v * g.ncell + c];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] += dtf * udot[v];
        if (fix_solution_mixed(g, u)) set_error(st, STAGE == 1 ? ERR_FIX_SOLUTION_1 : ERR_FIX_SOLUTION_2);
#pragma unroll
        for (int v = 0; v < 7; ++v) Udst[v * g.ncell + c] = u[v];
        if (STAGE == 2) {
            double q[7];
            q_from_u_mixed(q, u, g.gm1);
            fix_primitives_mixed(g, q, u[DEN]);
            inv_t = cfl_inv_t_mixed(g, q);
        }
    }
    if (STAGE == 2) block_max_to(inv_t, &st->maxinv, s.red);
}

template <int SOLVER>
static hipError_t launch_mixed_t(const HydroGeom& g, int stage, const double* Usrc, const double* Ubase,
                                 double* Udst, StepState* st, double dtf, hipStream_t stream) {
    dim3 grid(g.ntiles), block(NTHREADS);
    if (stage == 1)
        hipLaunchKernelGGL((hydro_stage_mixed_kernel<SOLVER, 1>), grid, block, 0, stream, g, Usrc, Ubase, Udst, st,
                           dtf);
    else
        hipLaunchKernelGGL((hydro_stage_mixed_kernel<SOLVER, 2>), grid, block, 0, stream, g, Usrc, Ubase, Udst, st,
                           dtf);
    return hipGetLastError();
}

hipError_t launch_hydro_stage_mixed(const HydroGeom& g, int solver, int stage, const double* Usrc,
                                    const double* Ubase, double* Udst, StepState* st, hipStream_t stream,
                                    double dtf) {
    switch (solver) {
    case SOLVER_HLL:  return launch_mixed_t<SOLVER_HLL>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    case SOLVER_HLLC: return launch_mixed_t<SOLVER_HLLC>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    default:          return launch_mixed_t<SOLVER_ROTATED>(g, stage, Usrc, Ubase, Udst, st, dtf, stream);
    }
}

}  // namespace tg
