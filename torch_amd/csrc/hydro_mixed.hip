// hydro_mixed.hip -- mixed-precision hydro stage kernel for CDNA4 (gfx950).
//
// The same stage as hydro.hip's FP64 kernel (Torch::hydroStep stage, see
// there for the reference map), with reconstruction and Riemann fluxes in
// float and the state, the flux accumulation, the primitive recovery, the
// floors and the CFL reduction in FP64.
//
// Layout for throughput: a workgroup of 256 threads (4 waves) owns a 16 x 16
// column of cells and marches it along z through MZ planes.  Each thread
// keeps the float primitives of its own column at planes k .. k+2 and the
// limited z slope of plane k in registers, so the z faces need neither LDS
// nor a reloaded z halo, and each z slope serves both z faces of its cell;
// plane k with its 2-deep x and y halos (256 + 128 cells, float [7][384],
// one plane per variable) goes through LDS.  Each thread limits the x and y
// slopes of its own cell there once (and 64 threads those of the 1-deep
// halo cells), so every slope serves both faces of its cell as in z; then
// the 544 x and y faces, which the 256 threads take as one flat list
// (consecutive lanes read consecutive x), are solved from the primitives
// and slopes of their two cells, and their fluxes replace the slopes in LDS
// once every face has read them.  The thread keeps the z flux of face
// k+1/2 in registers for the update and as face k-1/2 of the next plane.  Per cell that is 3.125 face
// solves (plus 1/MZ for the first z face of a column) and 1.5 cell loads, where
// 8^3 tiles with 2-deep halos in all three directions needed 3.375 and 2.5.
// After the faces each thread updates its cell of plane k in FP64 (fluxes
// accumulated in the reference's order x, y, z; += F_L, -= F_R), applies
// fixSolution, writes U, the float primitives of the output for the next
// stage and (corrector) its CFL term.  34 KB LDS, 128 VGPRs: four waves per
// SIMD, four workgroups per CU (neither the registers nor the LDS allow a
// fifth).
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


#ifndef TG_MTX
#define TG_MTX 16
#endif
constexpr int MTX = TG_MTX;               // x extent of a column
constexpr int MTY = 256 / MTX;            // y extent
constexpr int MCELLS = MTX * MTY;         // 256 cells per plane
constexpr int PW = MTX + 4, PH = MTY + 4; // plane with its 2-deep halos (corners unused)
constexpr int MHALO = 4 * (MTX + MTY);    // x and y halo cells per plane
constexpr int NXF = (MTX + 1) * MTY;      // x faces per plane
constexpr int NYF = MTX * (MTY + 1);      // y faces per plane
constexpr int MTHREADS = MCELLS;

constexpr int NSX = (MTX + 2) * MTY;      // cells with an x slope per plane (x in [-1, MTX])
constexpr int NSY = MTX * (MTY + 2);      // cells with a y slope per plane (y in [-1, MTY])
static_assert(NXF == NYF && NSX == NSY, "x and y lists of one length (square columns)");

struct MarchLDS {
    float p[7][PW * PH];    // plane k: cell (x, y) at (y + 2) PW + x + 2, x in [-2, MTX + 2), y in [-2, MTY + 2)
    // the limited slopes of plane k, read by the face solves; then (after a
    // barrier) the face fluxes, read by the update
    union {
        struct {
            float x[7][NSX];  // x slope of cell (x, y): y (MTX + 2) + x + 1
            float y[7][NSY];  // y slope of cell (x, y): (y + 1) MTX + x
        } s;
        struct {
            float x[7][NXF];  // x faces of plane k: y (MTX + 1) + f
            float y[7][NYF];  // y faces of plane k: f MTX + x
        } f;
    } sf;
    // a thread's flux of its first face, and of its third (the threads below
    // NXF + NYF - 2 MTHREADS), from the solve until the slopes are dead
    float park0[7][MTHREADS];
    float park2[7][NXF + NYF - 2 * MTHREADS];
    double red[MTHREADS / 64];
    unsigned int qmw[MTHREADS / 64][2];
};

__device__ __forceinline__ int mslot(int x, int y) { return (y + 2) * PW + x + 2; }

// Float primitives of global cell (x, y, z) (positions out of range are
// ghost cells: the first out-of-range dimension decides, as cell_source),
// scaled by 2^ks.  INNER: (x, y) lies inside the grid.
template <bool INNER, bool A32>
__device__ __forceinline__ void load_cell_f(const HydroGeom& g, const float* __restrict__ qin, int x, int y, int z,
                                            const int ks[7], float q[7]) {
    long long src;
    int mode = 0, mdim = 0;
    if (INNER && (unsigned)z < (unsigned)g.nz) {
        src = (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z);
    } else {
        const CellSrc cs = cell_source(g, x, y, z);
        src = cs.idx;
        mode = cs.mode;
        mdim = cs.mdim;
    }
#pragma unroll
    for (int v = 0; v < 7; ++v) q[v] = *soa_at<A32>(qin, g.ncell, v, src);
    if (mdim) apply_ghost_mode_f(q, mode, mdim - 1);
#pragma unroll
    for (int v = 0; v < 7; ++v) q[v] = __builtin_ldexpf(q[v], ks[v]);
}

// Halo cell h (0 .. MHALO - 1) of the plane: in-plane position relative to
// the column (x halos first, then y halos).
__device__ __forceinline__ void mhalo_pos(int h, int& x, int& y) {
    if (h < 4 * MTY) {
        const int l = h / MTY;
        x = l < 2 ? l - 2 : MTX + l - 2;
        y = h - l * MTY;
    } else {
        const int j = h - 4 * MTY;
        const int l = j / MTX;
        y = l < 2 ? l - 2 : MTY + l - 2;
        x = j - l * MTX;
    }
}

// Limited slope of cell b between its neighbours a and c (zero at first
// order): slope_flux adds it to b (times 1/2) on b's right face and
// subtracts it on its left face, so one evaluation serves both faces of b.
template <bool ALB>
__device__ __forceinline__ void cell_slope(const HydroGeom& g, const MixedScale& m, const float a[7], const float b[7],
                                           const float c[7], float s[7]) {
    if (g.spatial_order == 1) {
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            if constexpr (ALB) s[v] = albada_f(b[v] - a[v], c[v] - b[v], thr_of(m, v));
            else s[v] = limiter_f(g.limiter, b[v] - a[v], c[v] - b[v], thr_of(m, v));
        }
    } else {
#pragma unroll
        for (int v = 0; v < 7; ++v) s[v] = 0.f;
    }
}

// Face flux along axis DIM between cells b and c given their limited slopes
// sb and sc (cell_slope): face states b + sb/2 and c - sc/2.  With yswap
// (DIM 0 only) the face is a y face solved as an x face: the x and y
// velocity components of both face states are exchanged before the solve
// and of the flux after it (the Euler flux is covariant under that
// reflection, and the limiter threshold is the same for all velocity
// components).
template <int SOLVER, int DIM>
__device__ __forceinline__ bool slope_flux(const HydroGeom& g, const MixedScale& m, const float b[7], const float sb[7],
                                           const float c[7], const float sc[7], const GasConsts<float>& gc,
                                           float F[7], bool yswap = false) {
    float ql[7], qr[7];
    if (g.spatial_order == 1) {
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            ql[v] = b[v] + 0.5f * sb[v];
            qr[v] = c[v] - 0.5f * sc[v];
        }
    } else {
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            ql[v] = b[v];
            qr[v] = c[v];
        }
    }
    if (DIM == 0 && yswap) {
        const float lx = ql[VEL], rx = qr[VEL];
        ql[VEL] = ql[VEL + 1];
        qr[VEL] = qr[VEL + 1];
        ql[VEL + 1] = lx;
        qr[VEL + 1] = rx;
    }
    const bool bad = riemann_f<SOLVER, DIM>(F, ql, qr, gc, g.igm1f, m.nn_thr);
    if (DIM == 0 && yswap) {
        const float fx = F[VEL];
        F[VEL] = F[VEL + 1];
        F[VEL + 1] = fx;
    }
    return bad;
}

// The x and y slopes of plane k (cell_slope on the plane in LDS): each
// thread those of its own cell; threads below 4 MTX (MTX == MTY) the x slope
// of an x-halo cell (x = -1 or MTX) or the y slope of a y-halo cell (y = -1
// or MTY), which the boundary faces of the column read.
template <bool ALB>
__device__ __forceinline__ void march_slopes(const HydroGeom& g, MarchLDS& s, const MixedScale& m, int tid, int tx,
                                             int ty, const float w[7]) {
    float a[7], c[7], sl[7];
    const int o = mslot(tx, ty);
#pragma unroll
    for (int v = 0; v < 7; ++v) {
        a[v] = s.p[v][o - 1];
        c[v] = s.p[v][o + 1];
    }
    cell_slope<ALB>(g, m, a, w, c, sl);
#pragma unroll
    for (int v = 0; v < 7; ++v) s.sf.s.x[v][ty * (MTX + 2) + tx + 1] = sl[v];
#pragma unroll
    for (int v = 0; v < 7; ++v) {
        a[v] = s.p[v][o - PW];
        c[v] = s.p[v][o + PW];
    }
    cell_slope<ALB>(g, m, a, w, c, sl);
#pragma unroll
    for (int v = 0; v < 7; ++v) s.sf.s.y[v][(ty + 1) * MTX + tx] = sl[v];
    if (tid < 4 * MTX) {
        // halo cell h: x halos (x = -1, MTX; row h % MTY) then y halos
        const bool yh = tid >= 2 * MTY;
        const int h = yh ? tid - 2 * MTY : tid;
        const int side = h / MTY, r = h - side * MTY;
        const int hx = yh ? r : (side ? MTX : -1), hy = yh ? (side ? MTY : -1) : r;
        const int ho = mslot(hx, hy), st = yh ? PW : 1;
        float b[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            a[v] = s.p[v][ho - st];
            b[v] = s.p[v][ho];
            c[v] = s.p[v][ho + st];
        }
        cell_slope<ALB>(g, m, a, b, c, sl);
        float* dst = yh ? &s.sf.s.y[0][(hy + 1) * MTX + hx] : &s.sf.s.x[0][hy * (MTX + 2) + hx + 1];
#pragma unroll
        for (int v = 0; v < 7; ++v) dst[v * NSX] = sl[v];
    }
}

// One x face (item < NXF: row y = item / (MTX + 1), face f = item % (MTX + 1))
// or y face (j = item - NXF: face f = j / MTX, column x = j % MTX) of plane
// k from the primitives and limited slopes of its two cells: flux in F
// (zero for a face outside the grid).  Both kinds go through one code path
// (slope_flux's yswap), so the wave whose items straddle NXF does not run
// the solver once per axis.
template <int SOLVER, int STAGE>
__device__ __forceinline__ void march_xy_face(const HydroGeom& g, const MarchLDS& s, const MixedScale& m, int item,
                                              int ox, int oy, StepState* st, const GasConsts<float>& gc,
                                              float F[7]) {
    const bool yface = item >= NXF;
    int qb, qs, sb, ss;
    bool valid;
    if (!yface) {
        const int y = item / (MTX + 1), f = item - y * (MTX + 1);
        valid = (ox + f <= g.nx) && (oy + y < g.ny);
        qb = mslot(f - 1, y);
        qs = 1;
        sb = y * (MTX + 2) + f;
        ss = 1;
    } else {
        const int j = item - NXF;
        const int f = j / MTX, x = j - f * MTX;
        valid = (oy + f <= g.ny) && (ox + x < g.nx);
        qb = mslot(x, f - 1);
        qs = PW;
        sb = NSX * 7 + f * MTX + x;  // into the y slopes, which follow the x slopes
        ss = MTX;
    }
    bool bad = false;
    if (valid) {
        float b[7], c[7], sbv[7], scv[7];
        const float* sl = &s.sf.s.x[0][0];
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            b[v] = s.p[v][qb];
            c[v] = s.p[v][qb + qs];
            sbv[v] = sl[v * NSX + sb];
            scv[v] = sl[v * NSX + sb + ss];
        }
        bad = slope_flux<SOLVER, 0>(g, m, b, sbv, c, scv, gc, F, yface);
    } else {
#pragma unroll
        for (int v = 0; v < 7; ++v) F[v] = 0.f;
    }
    if (bad) set_error(st, STAGE == 1 ? ERR_RIEMANN_NAN_1 : ERR_RIEMANN_NAN_2);
}

__device__ __forceinline__ void store_xy_flux(MarchLDS& s, int item, const float F[7]) {
    float* dst = item >= NXF ? &s.sf.f.y[0][item - NXF] : &s.sf.f.x[0][item];
#pragma unroll
    for (int v = 0; v < 7; ++v) dst[v * NXF] = F[v];
}

template <int SOLVER, bool ALB, int STAGE, bool INNER, bool A32>
__device__ __forceinline__ void march_column(const HydroGeom& g, MarchLDS& s, const MixedScale& m, const int ks[7],
                                             int ox, int oy, int z0, int z1, const float* __restrict__ qin,
                                             const double* Ubase, double* Udst, float* __restrict__ qout,
                                             int qout_fix, StepState* st, double dtf_host, bool store_u,
                                             double& inv_t_max, unsigned int& mr, unsigned int& mp) {
    const int tid = threadIdx.x;
    const int tx = tid % MTX, ty = tid / MTX;
    const int gx = ox + tx, gy = oy + ty;
    const bool col = gx < g.nx && gy < g.ny;
    const GasConsts<float> gc = {g.gf, g.gm1f, g.g2f};
    const int ke[7] = {m.kr + m.kv, m.kp + m.kv, m.kr + m.kv, m.kr + m.kv, m.kp, m.kp, m.kp};
    const double dtf = stage_dtf(dtf_host, st, STAGE);
    // halo slot of this thread and its in-plane position
    int hx = 0, hy = 0;
    if (tid < MHALO) mhalo_pos(tid, hx, hy);

    // window of the own column: the limited z slope of plane k and the
    // primitives of planes k, k+1, k+2 (face k+1/2 needs the slopes of k and
    // k+1, and each slope serves the two z faces of its cell)
    float sk[7], w1[7], w2[7], w3[7], fz[7];
    {
        float a[7], w0[7], s0[7];
        load_cell_f<INNER, A32>(g, qin, gx, gy, z0 - 2, ks, a);
        load_cell_f<INNER, A32>(g, qin, gx, gy, z0 - 1, ks, w0);
        load_cell_f<INNER, A32>(g, qin, gx, gy, z0, ks, w1);
        load_cell_f<INNER, A32>(g, qin, gx, gy, z0 + 1, ks, w2);
        cell_slope<ALB>(g, m, a, w0, w1, s0);
        cell_slope<ALB>(g, m, w0, w1, w2, sk);
        bool bad = false;
        if (col) bad = slope_flux<SOLVER, 2>(g, m, w0, s0, w1, sk, gc, fz);   // face z0 - 1/2
        if (bad) set_error(st, STAGE == 1 ? ERR_RIEMANN_NAN_1 : ERR_RIEMANN_NAN_2);
        load_cell_f<INNER, A32>(g, qin, gx, gy, z0 + 2, ks, w3);
    }
    for (int k = z0; k < z1; ++k) {
        // plane k into LDS
#pragma unroll
        for (int v = 0; v < 7; ++v) s.p[v][mslot(tx, ty)] = w1[v];
        if (tid < MHALO) {
            float hq[7];
            load_cell_f<INNER, A32>(g, qin, ox + hx, oy + hy, k, ks, hq);
            const int hs = mslot(hx, hy);
#pragma unroll
            for (int v = 0; v < 7; ++v) s.p[v][hs] = hq[v];
        }
        if (STAGE == 1 && col) {
            // "hydro before" mark and the input's largest density and pressure
            const float rho = __builtin_ldexpf(w1[DEN], -ks[DEN]);
            if (rho != rho) set_error(st, ERR_CHECK_VALUES);
            mr = max(mr, max_bits_f(rho));
            mp = max(mp, max_bits_f(__builtin_ldexpf(w1[PRE], -ks[PRE])));
        }
        __syncthreads();
        // x and y slopes of plane k, then its x and y faces; each flux waits
        // (in LDS or, the last one solved, in registers) until every face has
        // read the slopes, whose LDS the fluxes then take over
        march_slopes<ALB>(g, s, m, tid, tx, ty, w1);
        __syncthreads();
        {
            constexpr int N3 = NXF + NYF - 2 * MTHREADS;
            const bool third = tid < N3;
            float F[7];
            if (third) {
                march_xy_face<SOLVER, STAGE>(g, s, m, tid + 2 * MTHREADS, ox, oy, st, gc, F);
#pragma unroll
                for (int v = 0; v < 7; ++v) s.park2[v][tid] = F[v];
            }
            march_xy_face<SOLVER, STAGE>(g, s, m, tid, ox, oy, st, gc, F);
#pragma unroll
            for (int v = 0; v < 7; ++v) s.park0[v][tid] = F[v];
            march_xy_face<SOLVER, STAGE>(g, s, m, tid + MTHREADS, ox, oy, st, gc, F);
            __syncthreads();
            store_xy_flux(s, tid + MTHREADS, F);
#pragma unroll
            for (int v = 0; v < 7; ++v) F[v] = s.park0[v][tid];
            store_xy_flux(s, tid, F);
            if (third) {
#pragma unroll
                for (int v = 0; v < 7; ++v) F[v] = s.park2[v][tid];
                store_xy_flux(s, tid + 2 * MTHREADS, F);
            }
        }
        // z face k + 1/2 from registers; the thread keeps its flux in registers
        // for the update (parked in LDS instead, it measured slower)
        float fzn[7];
        {
            float sn[7];
            cell_slope<ALB>(g, m, w1, w2, w3, sn);
            bool bad = false;
            if (col) bad = slope_flux<SOLVER, 2>(g, m, w1, sk, w2, sn, gc, fzn);
            else {
#pragma unroll
                for (int v = 0; v < 7; ++v) fzn[v] = 0.f;
            }
            if (bad) set_error(st, STAGE == 1 ? ERR_RIEMANN_NAN_1 : ERR_RIEMANN_NAN_2);
#pragma unroll
            for (int v = 0; v < 7; ++v) sk[v] = sn[v];
        }
        __syncthreads();
        if (col) {
            // += (A/V) F_L, -= (A/V) F_R per dimension (Hydro.cpp:117-137),
            // then the FP64 update of the cell (gravity is zero)
            double udot[7];
            const int fxl = ty * (MTX + 1) + tx, fyl = ty * MTX + tx;
#pragma unroll
            for (int v = 0; v < 7; ++v) {
                double u = 0.0;
                u += g.kflux * (double)s.sf.f.x[v][fxl];
                u -= g.kflux * (double)s.sf.f.x[v][fxl + 1];
                u += g.kflux * (double)s.sf.f.y[v][fyl];
                u -= g.kflux * (double)s.sf.f.y[v][fyl + MTX];
                u += g.kflux * (double)fz[v];
                u -= g.kflux * (double)fzn[v];
                udot[v] = ldexp(u, -ke[v]);
            }
            const long long c = (long long)gx + (long long)g.nx * ((long long)gy + (long long)g.ny * k);
            double u[7];
#pragma unroll
            for (int v = 0; v < 7; ++v) u[v] = *soa_at<A32>(Ubase, g.ncell, v, c);
#pragma unroll
            for (int v = 0; v < 7; ++v) u[v] += dtf * udot[v];
            double q[7];
            if (fix_solution_mixed_q(g, u, q)) set_error(st, STAGE == 1 ? ERR_FIX_SOLUTION_1 : ERR_FIX_SOLUTION_2);
            if (store_u) {
#pragma unroll
                for (int v = 0; v < 7; ++v) *soa_at<A32>(Udst, g.ncell, v, c) = u[v];
            }
            if (STAGE == 2) {
                fix_primitives_mixed(g, q, u[DEN]);
                inv_t_max = fmax(inv_t_max, cfl_inv_t_mixed(g, q));
            }
            store_qf<A32>(g, q, u, qout, c, qout_fix, STAGE == 2);
        }
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            fz[v] = fzn[v];
            w1[v] = w2[v];
            w2[v] = w3[v];
        }
        if (k + 1 < z1) load_cell_f<INNER, A32>(g, qin, gx, gy, k + 3, ks, w3);
    }
}

template <int SOLVER, bool ALB, int STAGE, bool A32>
// Four waves per SIMD (128 VGPRs; the 33 KB of LDS allow four workgroups
// per CU): about 20 registers spill, which measured faster than three waves
// per SIMD without spills (14.9 against 16.7 ms per 512^3 step).
__global__ __launch_bounds__(MTHREADS) __attribute__((amdgpu_waves_per_eu(4))) void hydro_stage_mixed_kernel(HydroGeom g, int mtx, int mty, int mtz, int mz,
                                                                     const float* __restrict__ qin, const double* Ubase,
                                                                     double* Udst, float* __restrict__ qout,
                                                                     int qout_fix, StepState* st, double dtf_host,
                                                                     int ustar_store) {
    __shared__ MarchLDS s;
    if (st->halted) return;
    // a two-stage corrector reads the predictor's float Q* only: the FP64 U*
    // is stored where something reads it (a single-stage sub-step's output,
    // a caller's stage, the last step of a run call, a step a print can
    // follow)
    const bool store_u = STAGE == 2 || ustar_store || st->keep_ustar;
    const int ntiles = mtx * mty * mtz;
    const int tile = tile_of_block(blockIdx.x, ntiles);
    const int bx = tile % mtx, by = (tile / mtx) % mty, bz = tile / (mtx * mty);
    const int ox = bx * MTX, oy = by * MTY, z0 = bz * mz;
    const int z1 = min(z0 + mz, g.nz);
    const MixedScale m = mixed_scale(st);
    const int ks[7] = {m.kr, m.kp, 0, 0, m.kv, m.kv, m.kv};
    double inv_t = 0.0;
    unsigned int mr = 0, mp = 0;
    const bool inner = ox >= 2 && oy >= 2 && ox + MTX + 2 <= g.nx && oy + MTY + 2 <= g.ny;
    if (inner)
        march_column<SOLVER, ALB, STAGE, true, A32>(g, s, m, ks, ox, oy, z0, z1, qin, Ubase, Udst, qout, qout_fix, st,
                                               dtf_host, store_u, inv_t, mr, mp);
    else
        march_column<SOLVER, ALB, STAGE, false, A32>(g, s, m, ks, ox, oy, z0, z1, qin, Ubase, Udst, qout, qout_fix, st,
                                                dtf_host, store_u, inv_t, mr, mp);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (STAGE == 1) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mr = max(mr, (unsigned int)__shfl_xor((int)mr, off, 64));
            mp = max(mp, (unsigned int)__shfl_xor((int)mp, off, 64));
        }
        if (lane == 0) {
            s.qmw[wave][0] = mr;
            s.qmw[wave][1] = mp;
        }
    } else {
        if (!(inv_t >= 0.0)) inv_t = 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) inv_t = fmax(inv_t, __shfl_xor(inv_t, off, 64));
        if (lane == 0) s.red[wave] = inv_t;
    }
    __syncthreads();
    if (tid == 0) {
        if (STAGE == 1) {
            unsigned int a = 0, b = 0;
            for (int w = 0; w < MTHREADS / 64; ++w) {
                a = max(a, s.qmw[w][0]);
                b = max(b, s.qmw[w][1]);
            }
            if (a) atomicMax(&st->qmax_next[0], a);
            if (b) atomicMax(&st->qmax_next[1], b);
        } else {
            double mx = s.red[0];
            for (int w = 1; w < MTHREADS / 64; ++w) mx = fmax(mx, s.red[w]);
            atomicMax(&st->maxinv, (unsigned long long)__double_as_longlong(mx));
        }
    }
}

// z planes per workgroup: 16 and 32 measured the same at 512^3, 16 ran 2 %
// faster at 256^3 (twice the workgroups).
#ifndef TG_MZ
#define TG_MZ 16
#endif
constexpr int MZ = TG_MZ;

template <int SOLVER, bool ALB>
static hipError_t launch_mixed_t(const HydroGeom& g, int stage, const float* qin, const double* Ubase, double* Udst,
                                 float* qout, int qout_fix, StepState* st, double dtf, int ustar, hipStream_t stream) {
    const int mtx = (g.nx + MTX - 1) / MTX, mty = (g.ny + MTY - 1) / MTY, mtz = (g.nz + MZ - 1) / MZ;
    dim3 grid(mtx * mty * mtz), block(MTHREADS);
    // 32-bit byte offsets within a variable while the FP64 arrays allow it
    // (ncell < 2^29, i.e. up to ~812^3 cells)
    const bool a32 = g.ncell * (long long)sizeof(double) < (1LL << 32) && !g.addr64;
    if (stage == 1) {
        if (a32)
            hipLaunchKernelGGL((hydro_stage_mixed_kernel<SOLVER, ALB, 1, true>), grid, block, 0, stream, g, mtx, mty,
                               mtz, MZ, qin, Ubase, Udst, qout, qout_fix, st, dtf, ustar);
        else
            hipLaunchKernelGGL((hydro_stage_mixed_kernel<SOLVER, ALB, 1, false>), grid, block, 0, stream, g, mtx, mty,
                               mtz, MZ, qin, Ubase, Udst, qout, qout_fix, st, dtf, ustar);
    } else {
        if (a32)
            hipLaunchKernelGGL((hydro_stage_mixed_kernel<SOLVER, ALB, 2, true>), grid, block, 0, stream, g, mtx, mty,
                               mtz, MZ, qin, Ubase, Udst, qout, qout_fix, st, dtf, 1);
        else
            hipLaunchKernelGGL((hydro_stage_mixed_kernel<SOLVER, ALB, 2, false>), grid, block, 0, stream, g, mtx, mty,
                               mtz, MZ, qin, Ubase, Udst, qout, qout_fix, st, dtf, 1);
    }
    return hipGetLastError();
}

// The van Albada instance carries no code of the other limiters: the
// generic limiter switch at every reconstruction site (14 per face solve,
// four inlined face solves) multiplies the kernel's code, and the 64 KB
// instruction cache is shared by two CUs.
template <int SOLVER>
static hipError_t launch_mixed_l(const HydroGeom& g, int stage, const float* qin, const double* Ubase, double* Udst,
                                 float* qout, int qout_fix, StepState* st, double dtf, int ustar, hipStream_t stream) {
    if (g.limiter == LIM_ALBADA)
        return launch_mixed_t<SOLVER, true>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, ustar, stream);
    return launch_mixed_t<SOLVER, false>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, ustar, stream);
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
                                    double dtf, int ustar) {
    if (stage == 1) hipLaunchKernelGGL(mixed_scale_rotate_kernel, dim3(1), dim3(1), 0, stream, st);
    switch (solver) {
    case SOLVER_HLL:
        return launch_mixed_l<SOLVER_HLL>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, ustar, stream);
    case SOLVER_HLLC:
        return launch_mixed_l<SOLVER_HLLC>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, ustar, stream);
    default:
        return launch_mixed_l<SOLVER_ROTATED>(g, stage, qin, Ubase, Udst, qout, qout_fix, st, dtf, ustar, stream);
    }
}

}  // namespace tg
