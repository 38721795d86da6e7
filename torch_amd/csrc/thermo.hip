// thermo.hip -- subcycled Henney heating/cooling source step on independent
// cells, FP64 (reference Thermodynamics::preTimeStepCalculations +
// integrate, src/Integrators/Thermodynamics.cpp:218-337), and the surrogate
// path that replaces the subcycle by the MLP of mlp.hip.
//
// The per-cell work is an explicit subcycle of nsteps = ceil(dt/dti) rate
// evaluations (:304-331); nsteps spans 1 .. ~1e8 across the cooling
// workload, so a lane-per-cell launch would leave most lanes idle behind the
// slowest lane of their wave.  The FP64 kernel is persistent instead: every
// lane owns one cell at a time and advances it by one subcycle per loop
// iteration; a lane whose cell is finished writes it out and takes the next
// work item from a global counter (one wave-aggregated atomic per refill).
// Work items may be an index list (the surrogate's fallback cells).
//
// Surrogate path (torchgpu_thermo_cells_surrogate), all on the device with
// no host round trip:
//   prepass   FP64 pre-step rates and subcycle count of every cell; cells
//             with at most `nexact` subcycles are integrated exactly here;
//             for the rest a coarse FP64 estimate Tc of the end temperature
//             (exponential-Euler steps of <= 2 % change) and the normalised
//             features are queued for the MLP (cells outside the training
//             envelope go to the fallback queue)
//   MLP       fused BF16-MFMA inference over the queue (mlp.hip); output 0
//             is log10(T_sub / Tc) / sigma(n_sub), a residual in units of
//             the coarse estimate's expected error, so the BF16 rounding of
//             the hidden activations scales with the residual; output 2 is
//             log10 of the estimated |error| of output 0 in the same units
//   finish    T_sub = Tc 10^(sigma y0) -> subcycled pressure -> RATE exactly
//             as :334; cells whose estimated relative T error sigma 10^y2
//             exceeds the gate are queued for the fallback
//   fallback  the persistent FP64 kernel over the fallback queue
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/torchgpu.h"
#include "hydro_device.hpp"
#include "hydro_math.hpp"
#include "mlp.hpp"
#include "star_device.hpp"
#include "thermo_math.hpp"
#include "torchgpu_internal.hpp"

namespace tg {

__device__ __forceinline__ double smin_d(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax_d(double a, double b) { return (a < b) ? b : a; }

struct CellState {
    double pre, heat, rate_pre, pressure, subT;
    double pre2temp, temp2pre, rate2dpre, dpre2rate, T_min, pf, T0;
    CoolingCell cc;
    long long remaining, nsub;
};

// Thermodynamics::preTimeStepCalculations for one non-wind cell followed by
// the set-up of integrate (:266-311).  Returns false when the cell is done
// (gate, no subcycling, or a single step).  pre_rate (optional): the
// cell's heating rate and pre-step RATE as this function computed them on
// the same state (pre_rate[pidx], pre_rate[pn + pidx]; thermo_timestep_kernel
// stores them), taken instead of evaluating them again.
__device__ __forceinline__ bool thermo_init(const ThermoConsts& t, const double* sx, const double* sy,
                                            const double* sy2, const double* in, long long stride, long long cell,
                                            CellState& s, unsigned int* err, const double* pre_rate = nullptr,
                                            long long pidx = 0, long long pn = 0) {
    const double den = in[0 * stride + cell], pre = in[1 * stride + cell], hii = in[2 * stride + cell];
    const double adv = in[3 * stride + cell], E = in[4 * stride + cell], T_min = in[5 * stride + cell];
    const double col_den = in[6 * stride + cell], F = in[7 * stride + cell], dt = in[8 * stride + cell];
    s.pre = pre;
    s.pressure = pre;
    s.heat = 0;
    s.rate_pre = 0;
    s.dpre2rate = 0;
    s.remaining = 0;
    s.nsub = 0;
    s.T0 = 0;
    if (adv < t.thermo_hii_switch) return false;  // :227-230, :275-280
    const double X = t.X;
    const double nH = X * den / t.hydrogen_mass;
    const double x = hii;
    const double mu_inv = X * (hii + 1.0) + (1.0 - X) * 0.25;
    const double T = (pre / den) * (1.0 / mu_inv) / t.spec_gas_const;  // Fluid::calcTemperature
    const double Av = 1.086 * t.dust_xsec * col_den;
    s.T0 = T;
    s.cc = cooling_cell(t, nH, x);
    double RATE;
    if (pre_rate) {
        s.heat = pre_rate[pidx];
        RATE = pre_rate[pn + pidx];
    } else {
        double rate = heating(t, nH, Av, F);
        s.heat = rate;
        rate = subtract_cooling(t, s.cc, sx, sy, sy2, rate, T);
        rate = soft_landing(rate, T, T_min);
        RATE = t.heating_amplification * rate;
    }
    s.rate_pre = RATE;
    if (!t.subcycling) return false;
    double dti = fabs(0.10 * E / RATE);
    s.pre2temp = 1.0 / (mu_inv * t.spec_gas_const * den);
    s.temp2pre = (mu_inv * t.spec_gas_const * den);
    s.rate2dpre = smin_d(dt, dti) * (t.gamma - 1.0);
    s.dpre2rate = 1.0 / s.rate2dpre;
    s.T_min = T_min;
    s.pf = smax_d(T_min * s.temp2pre, t.pfloor);
    double pressure = pre + RATE * s.rate2dpre;
    double subT = pressure * s.pre2temp;
    if (pressure < t.pfloor || subT < T_min) {
        subT = s.pf * s.pre2temp;
        pressure = s.pf;
    }
    s.pressure = pressure;
    s.subT = subT;
    if (dt > dti) {
        const double dtdti = dt / dti;
        if (!(dtdti < 2147483647.0)) {
            // the reference's int conversion (:307) is undefined here
            if (err) atomicOr(err, ERR_THERMO_NSTEPS);
            s.remaining = -1;
            return false;
        }
        int nsteps = dtdti - (int)dtdti > 0 ? (int)(dtdti + 1.0) : (int)(dtdti + 0.5);
        --nsteps;
        s.remaining = nsteps;
    }
    return s.remaining > 0;
}

// One subcycle, :314-330.
__device__ __forceinline__ void thermo_substep(const ThermoConsts& t, const double* sx, const double* sy,
                                               const double* sy2, CellState& s) {
    double r = subtract_cooling(t, s.cc, sx, sy, sy2, s.heat, s.subT);
    r = t.heating_amplification * soft_landing(r, s.subT, s.T_min);
    double pressure = s.pressure + r * s.rate2dpre;
    double subT = pressure * s.pre2temp;
    if (pressure < t.pfloor || subT < s.T_min) {
        subT = s.pf * s.pre2temp;
        pressure = s.pf;
    }
    s.pressure = pressure;
    s.subT = subT;
    s.nsub += 1;
    s.remaining -= 1;
}

__device__ __forceinline__ void thermo_finish(const CellState& s, double* out, long long stride, long long cell) {
    out[0 * stride + cell] = s.heat;
    out[1 * stride + cell] = s.rate_pre;
    // :334 (gate or no subcycling: RATE as preTimeStep left it)
    out[2 * stride + cell] = s.dpre2rate == 0 ? s.rate_pre : (s.pressure - s.pre) * s.dpre2rate;
    out[3 * stride + cell] = (double)s.nsub;
    out[4 * stride + cell] = s.pressure;
}

__device__ __forceinline__ long long shfl_ll(long long v, int src) {
    int lo = __shfl((int)(v & 0xffffffffll), src, 64);
    int hi = __shfl((int)(v >> 32), src, 64);
    return ((long long)hi << 32) | (unsigned int)lo;
}

// Wave-aggregated append: returns this lane's slot in list (only for lanes
// with pred true).
__device__ __forceinline__ long long wave_append(bool pred, unsigned long long* counter) {
    const unsigned long long mask = __ballot(pred);
    if (!mask) return -1;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    long long base = 0;
    if (lane == leader) base = (long long)atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = shfl_ll(base, leader);
    return pred ? base + __popcll(mask & ((1ull << lane) - 1ull)) : -1;
}

__device__ __forceinline__ void load_tables(const double* tables, double* sx, double* sy, double* sy2) {
    if (threadIdx.x < 26) {
        sx[threadIdx.x] = tables[threadIdx.x];
        sy[threadIdx.x] = tables[26 + threadIdx.x];
        sy2[threadIdx.x] = tables[52 + threadIdx.x];
    }
    __syncthreads();
}

// counters[0]: work counter; counters[1]: total subcycles.  Work items are
// 0..count-1 (count = *count_dev when given), cell = order ? order[i] : i.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void thermo_cells_kernel(ThermoConsts t, const double* __restrict__ tables,
                                                          long long stride, long long count,
                                                          const unsigned long long* __restrict__ count_dev,
                                                          const long long* __restrict__ order,
                                                          const double* __restrict__ in, double* __restrict__ out,
                                                          unsigned long long* __restrict__ counters,
                                                          unsigned int* __restrict__ err, unsigned char* path) {
    __shared__ double sx[26], sy[26], sy2[26];
    if (count_dev) count = (long long)*count_dev;
    if (count <= 0) return;  // an empty queue (the grid is sized for a full one)
    load_tables(tables, sx, sy, sy2);
    const int lane = threadIdx.x & 63;
    bool active = false, exhausted = false;
    long long cell = 0;
    unsigned long long nsub_total = 0;
    CellState s;
    while (true) {
        const unsigned long long need = __ballot(!active && !exhausted);
        if (need) {
            const int leader = __ffsll((long long)need) - 1;
            long long base = 0;
            if (lane == leader) base = (long long)atomicAdd(&counters[0], (unsigned long long)__popcll(need));
            base = shfl_ll(base, leader);
            if (!active && !exhausted) {
                const long long idx = base + __popcll(need & ((1ull << lane) - 1ull));
                if (idx < count) {
                    cell = order ? order[idx] : idx;
                    active = thermo_init(t, sx, sy, sy2, in, stride, cell, s, err);
                    if (path) path[cell] = order ? 2 : 0;
                    if (!active) thermo_finish(s, out, stride, cell);
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__any(active) && !__any(!exhausted)) break;
        if (active) {
            thermo_substep(t, sx, sy, sy2, s);
            if (s.remaining <= 0) {
                thermo_finish(s, out, stride, cell);
                nsub_total += (unsigned long long)s.nsub;
                active = false;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned int lo = __shfl_xor((unsigned int)(nsub_total & 0xffffffffull), off, 64);
        unsigned int hi = __shfl_xor((unsigned int)(nsub_total >> 32), off, 64);
        nsub_total += ((unsigned long long)hi << 32) | lo;
    }
    if (lane == 0) atomicAdd(&counters[1], nsub_total);
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------------------
// surrogate
// ---------------------------------------------------------------------------
struct SurrogateConsts {
    float mu[8], isd[8];
    double lo[7], hi[7];
    double gate;
    int nexact;
    int train;  // 1: queue every cell past nexact (no envelope), for training data
};

// Per-queued-cell data the finish step needs (FP64).
struct SurrCell {
    double T0, pre, temp2pre, pf, dpre2rate, T_min, heat, rate_pre;
    double Tc;     // coarse estimate of the subcycled end temperature
    double sigma;  // residual scale: the network predicts log10(T_sub / Tc) / sigma
};

// Scale (log10 units) of the coarse estimate's error for a subcycle of n
// steps: the explicit steps of short subcycles are coarse (large residuals),
// long ones resolve the evolution the coarse integration follows.  Dividing
// the residual by it keeps the network's output O(1), so the BF16 rounding
// of its hidden activations shrinks with the residual.
__device__ __forceinline__ double residual_scale(long long nsub) { return 0.01 / sqrt((double)nsub) + 2.0e-4; }

// Rate of one subcycle step at temperature T (Thermodynamics.cpp:314-320),
// as dp per subcycle.
__device__ __forceinline__ double substep_dp(const ThermoConsts& t, const double* sx, const double* sy,
                                             const double* sy2, const CellState& s, double T) {
    const double r = subtract_cooling(t, s.cc, sx, sy, sy2, s.heat, T);
    return t.heating_amplification * soft_landing(r, T, s.T_min) * s.rate2dpre;
}

// Coarse estimate of the temperature at the end of the subcycle: the
// remaining s.remaining explicit steps, each dp = rate(T) * rate2dpre, seen
// as the ODE dp/dk = rate(T(p)) * rate2dpre over k in [0, remaining] and
// integrated with exponential-Euler steps (linearised rate, Jacobian by a
// forward difference) that change p by at most ~2 % or span the rest, with
// the subcycle's floor after each step; at most COARSE_STEPS steps.  For
// long subcycles (the explicit steps resolve the evolution) this is close to
// the subcycle's own result; what remains is learned by the network.
constexpr int COARSE_STEPS = 512;
__device__ __forceinline__ double coarse_end_temperature(const ThermoConsts& t, const double* sx, const double* sy,
                                                         const double* sy2, const CellState& s) {
    double p = s.pressure;
    double kleft = (double)s.remaining;
    for (int it = 0; it < COARSE_STEPS && kleft > 0.0; ++it) {
        const double r0 = substep_dp(t, sx, sy, sy2, s, p * s.pre2temp);
        if (r0 == 0.0) break;
        const double eps = 1.0e-4 * p;
        const double r1 = substep_dp(t, sx, sy, sy2, s, (p + eps) * s.pre2temp);
        const double J = (r1 - r0) / eps;
        double dk = fmin(kleft, fmax(1.0, 0.02 * p / fabs(r0)));
        if (J > 0.0) dk = fmin(dk, fmax(1.0, 1.0 / J));
        const double x = J * dk;
        const double phi = fabs(x) < 1.0e-8 ? dk : expm1(x) / J;
        p += r0 * phi;
        if (p < t.pfloor || p * s.pre2temp < s.T_min) p = s.pf;
        kleft -= dk;
    }
    return p * s.pre2temp;
}

// Raw features u (code units): log10 rho, log10 T, x, log10 F_FUV (F_FUV
// floored at 1 code unit, so that an unlit cell, F_FUV = 0, has the finite
// feature 0; with the reference's unit scales 1 code unit is 1e-25
// photons cm^-2 s^-1), Av, log10 dt, and from the FP64 prepass log10 of the
// subcycle count and log10(Tc / T0) of the coarse end-temperature estimate.
// The training envelope applies to the first seven.
__device__ __forceinline__ void raw_features(const ThermoConsts& t, const double* in, long long stride, long long cell,
                                             const CellState& s, double Tc, double u[8]) {
    u[0] = log10(in[0 * stride + cell]);
    u[1] = log10(s.T0);
    u[2] = in[2 * stride + cell];
    u[3] = log10(fmax(in[7 * stride + cell], 1.0));
    u[4] = 1.086 * t.dust_xsec * in[6 * stride + cell];
    u[5] = log10(in[8 * stride + cell]);
    u[6] = log10((double)(s.remaining + 1));
    u[7] = log10(Tc / s.T0);
}

// The cooling operator's nine inputs of a grid cell (the column order of the
// cooling kernels: den, pre, hii, adv, E, T_min, col_den, F_FUV, dt) from the
// grid state: Q = QfromU(U) with fixPrimitives' clamps and floors, ADV = -1
// for the excluded cell and, with a star, its CausalWind cells, F_FUV from
// the star when no field is given.
struct GridSrc {
    HydroGeom g;
    const double* U;
    const double* T_min;     // NULL: T_min_default
    double T_min_default;
    const double* col_den;   // NULL: 0
    const double* F;         // NULL: the star's F_FUV, or 0
    double dt;               // >= 0, or DT_HALF / DT_FULL of the step in flight
    const StepState* st;
    long long excluded;
    StarGeom star;
};

__device__ __forceinline__ double grid_dt(const GridSrc& gs) {
    // DT_HALF / DT_FULL: the sub-step of the step in flight
    return gs.dt < 0.0 ? ((gs.dt == DT_HALF) ? 0.5 * gs.st->dt : gs.st->dt) : gs.dt;
}

__device__ __forceinline__ void gather_cell(const GridSrc& gs, long long c, double dt, double in[9],
                                            double u[7], double q[7]) {
    const HydroGeom& g = gs.g;
    const long long n = g.ncell;
#pragma unroll
    for (int v = 0; v < 7; ++v) u[v] = gs.U[v * n + c];
    QfromU(q, u, g.gm1);
    fix_primitives(g, q, u[DEN]);
    bool skip = c == gs.excluded;
    double f = gs.F ? gs.F[c] : 0.0;
    if (gs.star.on) {
        int ijk[3];
        cell_ijk32(gs.star, c, ijk);
        skip = skip || star_wind_cell(gs.star, ijk);
        if (!gs.F) f = star_fuv(gs.star, ijk);
    }
    in[0] = q[DEN];
    in[1] = q[PRE];
    in[2] = q[HII];
    in[3] = skip ? -1.0 : q[ADV];
    in[4] = u[PRE];
    in[5] = gs.T_min ? gs.T_min[c] : gs.T_min_default;
    in[6] = gs.col_den ? gs.col_den[c] : 0.0;
    in[7] = f;
    in[8] = dt;
}

// Thermodynamics::updateSourceTerms + Fluid::advSolution + fixSolution of
// one cell (apply_rate_fix_kernel's arithmetic): U[PRE] += dtf RATE, then
// fixSolution; with qf (a mixed-precision grid) the float Q of the result
// for the next hydro stage.  A halted run's cell keeps U but still gets
// its float Q.  U is not restrict: in the prepass it is the memory the
// cell's inputs were read from (GridSrc::U).
__device__ __forceinline__ void apply_cell(const HydroGeom& g, double u[7], double rate, double dtf, bool halted,
                                           double* U, float* __restrict__ qf, StepState* st,
                                           long long c) {
    if (!halted) {
        double udot[7] = {0, 0, 0, 0, 0, 0, 0};
        udot[PRE] += rate;
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] += dtf * udot[v];
        if (fix_solution(g, u)) set_error(st, ERR_THERMO_FIX);
#pragma unroll
        for (int v = 0; v < 7; ++v) U[v * g.ncell + c] = u[v];
    }
    if (qf) store_q_float(g, u, qf, c);
}

// A cell whose subcycle is longer than the exact limit: the coarse end
// temperature, its raw features and the training-envelope test, which
// route it to the network (to_mlp) or the FP64 fallback (to_fb).
__device__ __forceinline__ void route_active(const ThermoConsts& t, const double* sx, const double* sy,
                                             const double* sy2, const SurrogateConsts& sc, const double* cin,
                                             long long cstride, long long ccell, const CellState& s, double& Tc,
                                             double u[8], bool& to_mlp, bool& to_fb) {
    Tc = coarse_end_temperature(t, sx, sy, sy2, s);
    raw_features(t, cin, cstride, ccell, s, Tc, u);
    // the bounds are [min, max] of the training set's features, which were
    // stored as floats: compare the float-rounded features, so that a cell
    // equal to a training sample is in
    bool inside = true;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const float uf = (float)u[i];
        inside &= (uf >= (float)sc.lo[i]) && (uf <= (float)sc.hi[i]);
    }
    to_mlp = inside || sc.train;
    to_fb = !to_mlp;
}

// Queues a routed cell: network features and the state its finish needs,
// or the fallback list; path 1 / 2.  Every lane of the wave calls it (the
// appends are wave-aggregated).
__device__ __forceinline__ void queue_cell(bool to_mlp, bool to_fb, long long cell, const double u[8],
                                           const CellState& s, double Tc, const SurrogateConsts& sc,
                                           unsigned char* path, long long* __restrict__ mlp_list,
                                           float* __restrict__ feat, SurrCell* __restrict__ sdata,
                                           float* __restrict__ raw_out, long long* __restrict__ fb_list,
                                           unsigned long long* __restrict__ counters) {
    const long long slot = wave_append(to_mlp, &counters[2]);
    const long long fslot = wave_append(to_fb, &counters[3]);
    if (to_mlp) {
        mlp_list[slot] = cell;
#pragma unroll
        for (int i = 0; i < 8; ++i) feat[slot * 8 + i] = (float)(u[i] - (double)sc.mu[i]) * sc.isd[i];
        if (raw_out) {
#pragma unroll
            for (int i = 0; i < 8; ++i) raw_out[slot * 8 + i] = (float)u[i];
        }
        SurrCell d;
        d.T0 = s.T0;
        d.pre = s.pre;
        d.temp2pre = s.temp2pre;
        d.pf = s.pf;
        d.dpre2rate = s.dpre2rate;
        d.T_min = s.T_min;
        d.heat = s.heat;
        d.rate_pre = s.rate_pre;
        d.Tc = Tc;
        d.sigma = residual_scale(s.remaining + 1);
        sdata[slot] = d;
        if (path) path[cell] = 1;
    }
    if (to_fb) {
        fb_list[fslot] = cell;
        if (path) path[cell] = 2;
    }
}

// counters: [2] surrogate queue length, [3] fallback queue length.
// GRID: the inputs of each cell are formed from the grid state (gs, as
// thermo_gather_kernel) in registers, and written to `in` only for the
// cells queued for the network or the fallback, which read them later.
// A GRID cell finished here is also updated in the grid state (apply_cell
// into Uw / qf, dt from gs), and its `out` row is not written; the queued
// cells are updated by queue_apply_kernel once their rate is known (path
// marks them).
// Deferred cells of more remaining subcycles than this go to the front of
// the deferred list, the others to its back.
constexpr int COARSE_LONG = 64;

// DEFER (long_list given): every cell routed to the network or the fallback
// is queued in long_list instead, for surrogate_long_kernel, which estimates
// its coarse end temperature (up to COARSE_STEPS iterations of two rate
// evaluations) with every lane of a wave on such a cell; inline, one long
// cell in a wave keeps its other 63 lanes idle, and the estimate's registers
// would bound this kernel's occupancy.  Without DEFER (the training data)
// the routing is done here.  The GRID instance (which also forms the inputs
// and applies the update) runs at three waves per SIMD with a few spilled
// registers: measured faster than two without spills and four with ~50.
template <bool GRID, bool DEFER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GRID ? 3 : 4))) void surrogate_prepass_kernel(ThermoConsts t, const double* __restrict__ tables,
                                                               SurrogateConsts sc, long long n,
                                                               double* __restrict__ in, double* __restrict__ out,
                                                               unsigned char* path, long long* __restrict__ mlp_list,
                                                               float* __restrict__ feat, SurrCell* __restrict__ sdata,
                                                               float* __restrict__ raw_out,
                                                               long long* __restrict__ fb_list,
                                                               unsigned long long* __restrict__ counters,
                                                               unsigned int* __restrict__ err, GridSrc gs,
                                                               double* Uw, float* __restrict__ qf,
                                                               StepState* stw, long long* __restrict__ long_list,
                                                               const double* __restrict__ pre_rate) {
    __shared__ double sx[26], sy[26], sy2[26];
    load_tables(tables, sx, sy, sy2);
    const double gdt = GRID ? grid_dt(gs) : 0.0;
    const bool halted = GRID && gs.dt < 0.0 && gs.st->halted;
    const long long nloop = ((n + blockDim.x - 1) / blockDim.x) * blockDim.x;  // uniform trip count per wave
    for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x; cell - threadIdx.x < nloop;
         cell += (long long)gridDim.x * blockDim.x) {
        const bool valid = cell < n;
        bool to_mlp = false, to_fb = false, defer = false;
        CellState s;
        double u[8];
        double Tc = 0.0;
        double lin[9], lu[7], lq[7];
        const double* cin = in;  // this cell's inputs: cin[k * cstride + ccell]
        long long cstride = n, ccell = cell;
        if (GRID && valid) {
            gather_cell(gs, cell, gdt, lin, lu, lq);
            cin = lin;
            cstride = 1;
            ccell = 0;
        }
        if (valid) {
            bool active = thermo_init(t, sx, sy, sy2, cin, cstride, ccell, s, err, GRID ? pre_rate : nullptr, cell, n);
            if (active && s.remaining < sc.nexact) {
                while (s.remaining > 0) thermo_substep(t, sx, sy, sy2, s);
                active = false;
            }
            if (!active) {
                if (GRID) {
                    // the rate thermo_finish would store (Thermodynamics.cpp:334)
                    const double rate = s.dpre2rate == 0 ? s.rate_pre : (s.pressure - s.pre) * s.dpre2rate;
                    apply_cell(gs.g, lu, rate, gdt, halted, Uw, qf, stw, cell);
                } else {
                    thermo_finish(s, out, n, cell);
                }
                if (path) path[cell] = 0;
            } else {
                if (DEFER) defer = true;
                else route_active(t, sx, sy, sy2, sc, cin, cstride, ccell, s, Tc, u, to_mlp, to_fb);
                if (GRID) {
#pragma unroll
                    for (int k = 0; k < 9; ++k) in[k * n + cell] = lin[k];
                }
            }
        }
        if (DEFER) {
            // cells of many remaining subcycles from the front of long_list,
            // the others from its back, so that the former share waves
            const bool lng = defer && s.remaining > COARSE_LONG;
            const long long fslot = wave_append(lng, &counters[4]);
            const long long bslot = wave_append(defer && !lng, &counters[5]);
            if (lng) long_list[fslot] = cell;
            else if (defer) long_list[n - 1 - bslot] = cell;
        } else {
            queue_cell(to_mlp, to_fb, cell, u, s, Tc, sc, path, mlp_list, feat, sdata, raw_out, fb_list, counters);
        }
    }
}

// The deferred cells of the prepass (long_list: counters[4] from its front,
// counters[5] from its back, of n slots): the same routing from their
// inputs in `in` (column k of cell c at k n + c; the grid prepass writes
// them there), after thermo_init again.  One wave
// per workgroup, so that a few hundred waves of such cells spread over all
// CUs, and two waves per SIMD, which leave the coarse loop's registers
// unspilled.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void surrogate_long_kernel(
    ThermoConsts t, const double* __restrict__ tables, SurrogateConsts sc, long long n, const double* __restrict__ in,
    unsigned char* path, long long* __restrict__ mlp_list, float* __restrict__ feat, SurrCell* __restrict__ sdata,
    long long* __restrict__ fb_list, unsigned long long* __restrict__ counters,
    const long long* __restrict__ long_list, unsigned int* __restrict__ err) {
    __shared__ double sx[26], sy[26], sy2[26];
    const long long nfront = (long long)counters[4];
    const long long count = nfront + (long long)counters[5];
    if ((long long)blockIdx.x * blockDim.x >= count) return;  // uniform per block, before the table load's barrier
    load_tables(tables, sx, sy, sy2);
    const long long nloop = ((count + blockDim.x - 1) / blockDim.x) * blockDim.x;  // uniform trip count per wave
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i - threadIdx.x < nloop;
         i += (long long)gridDim.x * blockDim.x) {
        bool to_mlp = false, to_fb = false;
        CellState s;
        double u[8];
        double Tc = 0.0;
        long long cell = 0;
        if (i < count) {
            cell = i < nfront ? long_list[i] : long_list[n - 1 - (i - nfront)];
            (void)thermo_init(t, sx, sy, sy2, in, n, cell, s, err);
            route_active(t, sx, sy, sy2, sc, in, n, cell, s, Tc, u, to_mlp, to_fb);
        }
        queue_cell(to_mlp, to_fb, cell, u, s, Tc, sc, path, mlp_list, feat, sdata, nullptr, fb_list, counters);
    }
}

// The grid update (apply_cell) of the cells the surrogate step queued
// (path 1: the network's rate; path 2: the FP64 fallback's), after both
// have run; the prepass updated the rest.  cerr: the cooling operator's
// error word; a subcycle count beyond the int range is moved into the step
// state (reported at this step, as the fixSolution errors) and cleared.
// counters: the surrogate step's, [2] network and [3] fallback queue
// lengths; with both queues empty there is no cell to update.
__global__ void queue_apply_kernel(GridSrc gs, const double* __restrict__ out, const unsigned char* __restrict__ path,
                                   double* __restrict__ U, float* __restrict__ qf, StepState* st, unsigned int* cerr,
                                   const unsigned long long* __restrict__ counters) {
    const bool halted = gs.dt < 0.0 && gs.st->halted;
    if (!halted && cerr && blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned int e = *cerr;
        if (e & ERR_THERMO_NSTEPS) {
            set_error(st, ERR_THERMO_NSTEPS);
            *cerr = 0u;
        }
    }
    if (counters[2] == 0 && counters[3] == 0) return;
    const double dtf = grid_dt(gs);
    const long long n = gs.g.ncell;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        if (path[c] == 0) continue;
        double u[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * n + c];
        apply_cell(gs.g, u, out[2 * n + c], dtf, halted, U, qf, st, c);
    }
}

// y (fp32 [count][3]) -> subcycled pressure and RATE; gated cells -> fallback.
__global__ void surrogate_finish_kernel(ThermoConsts t, SurrogateConsts sc, long long stride,
                                        const unsigned long long* __restrict__ counters_in,
                                        const long long* __restrict__ mlp_list, const SurrCell* __restrict__ sdata,
                                        const float* __restrict__ y, double* __restrict__ out, unsigned char* path,
                                        long long* __restrict__ fb_list, unsigned long long* __restrict__ counters) {
    const long long count = (long long)counters_in[2];
    const long long nloop = ((count + blockDim.x - 1) / blockDim.x) * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i - threadIdx.x < nloop;
         i += (long long)gridDim.x * blockDim.x) {
        bool to_fb = false;
        long long cell = 0;
        if (i < count) {
            cell = mlp_list[i];
            const SurrCell d = sdata[i];
            const double y0 = (double)y[i * 3 + 0];
            // output 2 estimates log10 |error of output 0| (scaled units, ln 10 included)
            const double est = d.sigma * exp10((double)y[i * 3 + 2]);
            if (!(est <= sc.gate) || !isfinite(y0)) {
                to_fb = true;
            } else {
                const double Tsub = d.Tc * exp10(d.sigma * y0);
                double p = Tsub * d.temp2pre;
                if (p < t.pfloor || Tsub < d.T_min) p = d.pf;
                out[0 * stride + cell] = d.heat;
                out[1 * stride + cell] = d.rate_pre;
                out[2 * stride + cell] = (p - d.pre) * d.dpre2rate;
                out[3 * stride + cell] = 0.0;
                out[4 * stride + cell] = p;
            }
        }
        const long long fslot = wave_append(to_fb, &counters[3]);
        if (to_fb) {
            fb_list[fslot] = cell;
            if (path) path[cell] = 2;
        }
    }
}

// Training targets of the queued cells from the FP64 integrator's output:
// t0 = (log10(p_sub * pre2temp) - log10(Tc)) / sigma with pre2temp =
// 1/temp2pre.
__global__ void surrogate_targets_kernel(const unsigned long long* __restrict__ counters, long long stride,
                                         const long long* __restrict__ mlp_list, const SurrCell* __restrict__ sdata,
                                         const double* __restrict__ out, float* __restrict__ target) {
    const long long count = (long long)counters[2];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
        const SurrCell d = sdata[i];
        const double p = out[4 * stride + mlp_list[i]];
        const double Tsub = p * (1.0 / d.temp2pre);
        target[i * 3 + 0] = (float)((log10(Tsub) - log10(d.Tc)) / d.sigma);
        target[i * 3 + 1] = 0.f;
        target[i * 3 + 2] = 0.f;
    }
}

// ---------------------------------------------------------------------------
// Thermodynamics on the grid state (operator-split step)
// ---------------------------------------------------------------------------
// Per-cell inputs of the cooling kernels from the grid: Q =
// fixPrimitives(QfromU(U)) (Fluid.cpp:138-165), E = U[PRE], T_min, COL_DEN,
// F_FUV (arrays or defaults) and dt.  The excluded cell gets ADV = -1 so
// that the gate makes it a no-op; with a star, so do its CausalWind cells
// (Fluid::initialiseGrid, Fluid.cpp:41-54), which the reference's
// preTimeStepCalculations skips, and F_FUV is the star's when no array is
// given (Thermodynamics.cpp:237-243).
__global__ void thermo_gather_kernel(GridSrc gs, double* __restrict__ in9) {
    const long long n = gs.g.ncell;
    const double dt = grid_dt(gs);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        double in[9], u[7], q[7];
        gather_cell(gs, c, dt, in, u, q);
#pragma unroll
        for (int k = 0; k < 9; ++k) in9[k * n + c] = in[k];
    }
}

// Thermodynamics::calculateTimeStep (Thermodynamics.cpp:463-476): min over
// cells with RATE != 0 of |f U[PRE] / RATE|, f = 1 when subcycling else 0.1,
// from the pre-step RATE; reduced as the bit pattern of a non-negative double.
// GRID: the inputs from the grid state (gs) in registers, no gathered array;
// with cfl_st, also the hydro CFL maximum of the same state into
// cfl_st->maxinv (cfl_kernel's: the same fixPrimitives(QfromU(U)) per cell,
// and a maximum, which any order of reduction gives bitwise); with rate_out,
// each cell's heating rate and pre-step RATE ([2][n]) for a cooling sub-step
// on the same state (thermo_init's pre_rate).
template <bool GRID>
__global__ __launch_bounds__(256) void thermo_timestep_kernel(ThermoConsts t, const double* __restrict__ tables,
                                                             long long n, const double* __restrict__ in9,
                                                             unsigned long long* __restrict__ dtmin_bits,
                                                             unsigned int* __restrict__ err, GridSrc gs,
                                                             StepState* cfl_st, double* __restrict__ rate_out) {
    __shared__ double sx[26], sy[26], sy2[26];
    __shared__ double red[4], redc[4];
    load_tables(tables, sx, sy, sy2);
    double m = __longlong_as_double(0x7fefffffffffffffll);
    double mc = 0.0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        CellState s;
        ThermoConsts t1 = t;
        t1.subcycling = 0;  // pre-step rate only
        double lin[9], lu[7], lq[7];
        const double* cin = in9;
        long long cstride = n, ccell = c;
        if (GRID) {
            gather_cell(gs, c, 0.0, lin, lu, lq);
            cin = lin;
            cstride = 1;
            ccell = 0;
            if (cfl_st) {
                const double it = cfl_inv_t(gs.g, lq);
                if (it >= 0.0) mc = fmax(mc, it);
            }
        }
        (void)thermo_init(t1, sx, sy, sy2, cin, cstride, ccell, s, err);
        if (GRID && rate_out) {
            rate_out[c] = s.heat;
            rate_out[n + c] = s.rate_pre;
        }
        if (s.rate_pre != 0) {
            const double frac = t.subcycling ? 1.0 : 0.1;
            const double dti = fabs(frac * cin[4 * cstride + ccell] / s.rate_pre);
            if (dti < m) m = dti;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m = fmin(m, __shfl_xor(m, off, 64));
        mc = fmax(mc, __shfl_xor(mc, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = m;
        redc[threadIdx.x >> 6] = mc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
        if (b >= 0.0) atomicMin(dtmin_bits, (unsigned long long)__double_as_longlong(b));
        if (GRID && cfl_st) {
            const double bc = fmax(fmax(redc[0], redc[1]), fmax(redc[2], redc[3]));
            atomicMax(&cfl_st->maxinv, (unsigned long long)__double_as_longlong(bc));
        }
    }
}

__global__ void set_u64_kernel(unsigned long long* p, unsigned long long v) { *p = v; }

__global__ void zero_f64_kernel(double* p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// GridCell::H, the heating arrays Torch::run prints at checkpoints.  A grid
// that tracks them keeps TOT, RHII and EUVH between calls (heat [3][n]):
// Thermodynamics::integrate sets TOT to the post-subcycle RATE of each
// non-wind cell, or zeroes all of H below the HII switch
// (Thermodynamics.cpp:275-280,334-335); Radiation::preTimeStepCalculations
// sets RHII and EUVH (Radiation.cpp:278-279); fillHeatingArrays recomputes
// the other eight from Q and adds RHII + EUVH to TOT (:419-461).
// ---------------------------------------------------------------------------
__global__ void heat_integrate_kernel(long long n, const double* __restrict__ in9, const double* __restrict__ out5,
                                      double dt, const StepState* st, double hii_switch, double* __restrict__ heat) {
    if (dt < 0.0 && st->halted) return;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        const double adv = in9[3 * n + c];
        if (adv == -1.0) continue;  // a wind cell (the gather's mark): never integrated
        if (adv < hii_switch) {
            heat[HEAT_TOT * n + c] = 0.0;
            heat[HEAT_RHII * n + c] = 0.0;
            heat[HEAT_EUVH * n + c] = 0.0;
        } else {
            heat[HEAT_TOT * n + c] = out5[2 * n + c];
        }
    }
}

// Thermodynamics::fillHeatingArrays on the primitives Q [7][n] of the
// printed state; col_den: COL_DEN per cell (the star's ray trace or the
// caller's array), or null for zero; F: the caller's F_FUV or null (the
// star's flux when a star is on, else zero).  H [11][n] in HID order.
__global__ __launch_bounds__(256) void fill_heating_kernel(HydroGeom g, ThermoConsts t, const double* __restrict__ tables,
                                                           const double* __restrict__ Q,
                                                           const double* __restrict__ col_den,
                                                           const double* __restrict__ F, StarGeom star,
                                                           long long excluded, double* __restrict__ heat,
                                                           double* __restrict__ H) {
    __shared__ double sx[26], sy[26], sy2[26];
    load_tables(tables, sx, sy, sy2);
    const long long n = g.ncell;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        const int ijk[3] = {(int)(c % g.nx), (int)((c / g.nx) % g.ny), (int)(c / ((long long)g.nx * g.ny))};
        bool wind = c == excluded;
        if (star.on) wind = wind || star_wind_cell(star, ijk);
        double h[11];
#pragma unroll
        for (int i = 0; i < 11; ++i) h[i] = 0.0;
        if (!wind) {
            if (Q[ADV * n + c] < t.thermo_hii_switch) {
                heat[HEAT_TOT * n + c] = 0.0;
                heat[HEAT_RHII * n + c] = 0.0;
                heat[HEAT_EUVH * n + c] = 0.0;
            } else {
                const double den = Q[DEN * n + c], pre = Q[PRE * n + c], x = Q[HII * n + c];
                const double nH = t.X * den / t.hydrogen_mass;
                // GridCell::temperature (GridCell.cpp:158-161)
                const double mu_inv = t.X * (x + 1.0) + (1.0 - t.X) * 0.25;
                const double T = (pre / den) * (1.0 / mu_inv) / t.spec_gas_const;
                double f = F ? F[c] : 0.0;
                if (star.on && !F) f = star_fuv(star, ijk);
                const double tau = col_den ? col_den[c] : 0.0;
                const double Av = 1.086 * t.dust_xsec * tau;
                heating_terms(t, sx, sy, sy2, nH, x, T, Av, f, h);
                const double rhii = heat[HEAT_RHII * n + c], euvh = heat[HEAT_EUVH * n + c];
                const double tot = heat[HEAT_TOT * n + c] + (rhii + euvh);
                heat[HEAT_TOT * n + c] = tot;
                h[2] = rhii;
                h[6] = euvh;
                h[10] = tot;
            }
        }
#pragma unroll
        for (int i = 0; i < 11; ++i) H[i * n + c] = h[i];
    }
}

// ---------------------------------------------------------------------------
// Longest-first work order for the persistent FP64 kernel.  A cell's cost is
// its subcycle count, known from the pre-step rate (thermo_init), and spans
// 1 .. ~1e8; handed out in index order, a long cell that starts late ends
// the launch long after the rest.  Cells are bucketed by floor(log2(count))
// and the kernel takes the buckets longest first (LPT scheduling, within a
// factor of two of the exact order), so the launch ends near
// max(longest cell, total work / throughput).  Results do not depend on the
// order: every cell is integrated independently.
// ---------------------------------------------------------------------------
constexpr int LPT_NB = 64;

__global__ __launch_bounds__(256) void lpt_bucket_kernel(ThermoConsts t, const double* __restrict__ tables,
                                                        long long stride, long long count,
                                                        const unsigned long long* __restrict__ count_dev,
                                                        const long long* __restrict__ list,
                                                        const double* __restrict__ in,
                                                        unsigned char* __restrict__ bucket,
                                                        unsigned long long* __restrict__ hist,
                                                        unsigned int* __restrict__ err) {
    __shared__ double sx[26], sy[26], sy2[26];
    __shared__ unsigned int h[LPT_NB];
    if (threadIdx.x < LPT_NB) h[threadIdx.x] = 0;
    load_tables(tables, sx, sy, sy2);
    if (count_dev) count = (long long)*count_dev;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x) {
        const long long cell = list ? list[i] : i;
        CellState s;
        thermo_init(t, sx, sy, sy2, in, stride, cell, s, err);
        const int b = s.remaining > 0 ? 64 - __clzll(s.remaining) : 0;
        bucket[i] = (unsigned char)b;
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < LPT_NB && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// Subcycle count of each cell (the explicit steps integrate() takes after
// the first, fixed by the pre-step rate): 0 for a cell that is not
// subcycled, -1 where dt / dti is beyond the int range.  Nothing is
// integrated and no error is raised.
__global__ __launch_bounds__(256) void subcycle_count_kernel(ThermoConsts t, const double* __restrict__ tables,
                                                            long long n, const double* __restrict__ in,
                                                            long long* __restrict__ count) {
    __shared__ double sx[26], sy[26], sy2[26];
    load_tables(tables, sx, sy, sy2);
    for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x; cell < n;
         cell += (long long)gridDim.x * blockDim.x) {
        CellState s;
        thermo_init(t, sx, sy, sy2, in, n, cell, s, nullptr);
        count[cell] = s.remaining;
    }
}

// cursor[b] = number of items in buckets above b (longest first)
__global__ void lpt_offsets_kernel(const unsigned long long* __restrict__ hist, unsigned long long* __restrict__ cursor) {
    if (threadIdx.x != 0) return;
    unsigned long long off = 0;
    for (int b = LPT_NB - 1; b >= 0; --b) {
        cursor[b] = off;
        off += hist[b];
    }
}

__global__ __launch_bounds__(256) void lpt_scatter_kernel(long long count, const unsigned long long* __restrict__ count_dev,
                                                         const long long* __restrict__ list,
                                                         const unsigned char* __restrict__ bucket,
                                                         unsigned long long* __restrict__ cursor,
                                                         long long* __restrict__ order) {
    __shared__ unsigned int h[LPT_NB];
    __shared__ unsigned long long base[LPT_NB];
    if (count_dev) count = (long long)*count_dev;
    const int tid = threadIdx.x;
    // the trip count is uniform over the block (barriers inside)
    for (long long i0 = (long long)blockIdx.x * blockDim.x; i0 < count; i0 += (long long)gridDim.x * blockDim.x) {
        if (tid < LPT_NB) h[tid] = 0;
        __syncthreads();
        const long long i = i0 + tid;
        int b = -1;
        unsigned int rank = 0;
        if (i < count) {
            b = bucket[i];
            rank = atomicAdd(&h[b], 1u);
        }
        __syncthreads();
        if (tid < LPT_NB && h[tid]) base[tid] = atomicAdd(&cursor[tid], (unsigned long long)h[tid]);
        __syncthreads();
        if (b >= 0) order[base[b] + rank] = list ? list[i] : i;
        __syncthreads();
    }
}

}  // namespace tg

struct torchgpu_cooling {
    torchgpu_thermo_params p;
    tg::ThermoConsts t;
    hipStream_t stream = nullptr;
    double* tables = nullptr;                 // spline x, y, y2 (78 doubles)
    unsigned long long* counters = nullptr;   // [0] work, [1] subcycles, [2] surrogate queue, [3] fallback queue,
                                              // [4], [5] cells the surrogate prepass deferred (long, short)
    unsigned int* err = nullptr;
    char msg[256] = {0};
    int nsm = 256;
    // surrogate
    torchgpu_mlp* mlp = nullptr;
    tg::SurrogateConsts sc{};
    bool has_surrogate = false;
    long long cap = 0;                        // capacity of the queues below
    long long* mlp_list = nullptr;
    long long* fb_list = nullptr;
    long long* long_list = nullptr;           // the surrogate prepass's deferred cells
    float* feat = nullptr;
    float* y = nullptr;
    tg::SurrCell* sdata = nullptr;
    // longest-first order of the FP64 kernel's work items
    long long lpt_cap = 0;
    long long* lpt_order = nullptr;
    unsigned char* lpt_bucket = nullptr;
    unsigned long long* lpt_hist = nullptr;   // [0, 64) histogram, [64, 128) cursors
    // grid-operator scratch
    long long grid_cap = 0;
    double* in9 = nullptr;
    double* out5 = nullptr;
    unsigned char* gpath = nullptr;  // per grid cell: the surrogate step's path (0 exact, 1 network, 2 fallback)
    double* rcache = nullptr;        // [2][n]: heating rate and pre-step RATE of the thermo time step's state
    long long rcache_cap = 0;        // (allocated by the first time step asked to keep them)
    unsigned long long* dtbits = nullptr;
    // star (torchgpu_cooling_set_star): geometry and the DCOL_DEN of the
    // ray trace (COL_DEN goes straight into in9)
    tg::StarGeom star{};
    // Thermodynamics::initialiseMinTempField's per-cell T_min, taken when a
    // grid call passes no T_min (else the constant m_T_min = 100 K)
    double* tmin = nullptr;
    long long tmin_n = 0;
    double* dcol = nullptr;
    double* col = nullptr;   // COL_DEN of the last ray trace
    long long dcol_cap = 0;
    double* xface = nullptr;   // the ray trace's x-face buffers
    long long xface_cap = 0;
    int device = 0;            // the HIP device current at create
};

// ---------------------------------------------------------------------------
// Fixed scratch of torchgpu_thermo_cells_noalloc: the longest-first work
// order of at most NOALLOC_MAX_CELLS cells (histogram and cursors, per-cell
// buckets, order), in one device array per GPU that lives until the process
// ends.  Host code never passes `scratch` itself to a kernel (on the host the
// name is a stand-in, not a device address): each device's address comes
// from hipGetSymbolAddress, taken once per device at set-up.
// ---------------------------------------------------------------------------
constexpr long long NOALLOC_MAX_CELLS = 1LL << 20;
constexpr size_t NOALLOC_HIST_BYTES = 2 * tg::LPT_NB * sizeof(unsigned long long);   // 1 KiB
constexpr size_t NOALLOC_ORDER_BYTES = NOALLOC_MAX_CELLS * sizeof(long long);         // 8 MiB
constexpr size_t NOALLOC_BYTES = NOALLOC_HIST_BYTES + NOALLOC_ORDER_BYTES + NOALLOC_MAX_CELLS;
static_assert(NOALLOC_HIST_BYTES % 256 == 0 && NOALLOC_ORDER_BYTES % 256 == 0, "scratch sections stay aligned");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// the MLP handle is defined in mlp_api.cpp; the cooling operator only needs
// its device weights
extern "C" int torchgpu_mlp_device_weights(torchgpu_mlp* m, const uint16_t** w16, const float** b);

namespace {

#define HIP_OK(call)                                 \
    do {                                             \
        hipError_t e_ = (call);                      \
        if (e_ != hipSuccess) {                      \
            if (c) snprintf(c->msg, sizeof c->msg, "%s: %s", #call, hipGetErrorString(e_)); \
            return TORCHGPU_ERR_HIP;                 \
        }                                            \
    } while (0)

tg::ThermoConsts to_consts(const torchgpu_thermo_params& p) {
    tg::ThermoConsts t{};
    t.thermo_hii_switch = p.thermo_hii_switch;
    t.heating_amplification = p.heating_amplification;
    t.X = p.mass_fraction_h;
    t.gamma = p.gamma;
    t.pfloor = p.pfloor;
    t.hydrogen_mass = p.hydrogen_mass;
    t.spec_gas_const = p.spec_gas_const;
    t.dust_xsec = p.dust_xsec;
    t.pi = p.pi;
    t.z0 = p.z0; t.T1 = p.T1; t.T2 = p.T2; t.T3 = p.T3; t.T4 = p.T4;
    t.imlc = p.imlc; t.nmlc = p.nmlc; t.ciec_minT = p.ciec_minT; t.ciec = p.ciec; t.cxhi_damp = p.cxhi_damp;
    t.n0 = p.n0; t.nmc = p.nmc;
    t.fuvh_a = p.fuvh_a; t.fuvh_b = p.fuvh_b; t.fuvh_c = p.fuvh_c; t.irh_a = p.irh_a; t.irh_b = p.irh_b;
    t.crh = p.crh;
    t.spl_min_slope = p.spl_min_slope;
    t.spl_max_slope = p.spl_max_slope;
    t.subcycling = p.subcycling;
    return t;
}

int zero_counters(torchgpu_cooling* c) {
    hipLaunchKernelGGL(tg::zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, c->counters, 6);
    HIP_OK(hipGetLastError());
    return TORCHGPU_OK;
}

int launch_fp64(torchgpu_cooling* c, long long stride, long long count, const unsigned long long* count_dev,
                const long long* order, const double* in, double* out, unsigned char* path) {
    long long blocks = (long long)c->nsm * 8;  // 32 waves per CU
    if (!count_dev && blocks * 256 > count) blocks = (count + 255) / 256;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(tg::thermo_cells_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, c->t, c->tables,
                       stride, count, count_dev, order, in, out, c->counters, c->err, path);
    HIP_OK(hipGetLastError());
    return TORCHGPU_OK;
}

// Work order (longest cells first) of `count` items (count_dev: on the
// device) that are cells list[i] (list == NULL: i), into `order` [count],
// with the per-item buckets in `bucket` [count] and the histogram and
// cursors in `hist` [2 LPT_NB] (cleared here by a kernel).
int lpt_order_into(torchgpu_cooling* c, long long stride, long long count, const unsigned long long* count_dev,
                   const long long* list, const double* in, long long* order, unsigned char* bucket,
                   unsigned long long* hist) {
    hipLaunchKernelGGL(tg::zero_u64_kernel, dim3(1), dim3(2 * tg::LPT_NB), 0, c->stream, hist, 2 * tg::LPT_NB);
    HIP_OK(hipGetLastError());
    long long blocks = (count + 255) / 256;
    if (blocks > (long long)c->nsm * 16) blocks = (long long)c->nsm * 16;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(tg::lpt_bucket_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, c->t, c->tables,
                       stride, count, count_dev, list, in, bucket, hist, c->err);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(tg::lpt_offsets_kernel, dim3(1), dim3(64), 0, c->stream, hist, hist + tg::LPT_NB);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(tg::lpt_scatter_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, count, count_dev,
                       list, bucket, hist + tg::LPT_NB, order);
    HIP_OK(hipGetLastError());
    return TORCHGPU_OK;
}

// lpt_order_into on the handle's own buffers (grown as needed); returns
// c->lpt_order.
int lpt_order(torchgpu_cooling* c, long long stride, long long count, const unsigned long long* count_dev,
              const long long* list, const double* in, const long long** order) {
    if (count > c->lpt_cap) {
        (void)hipStreamSynchronize(c->stream);
        if (c->lpt_order) (void)hipFree(c->lpt_order);
        if (c->lpt_bucket) (void)hipFree(c->lpt_bucket);
        c->lpt_order = nullptr;
        c->lpt_bucket = nullptr;
        c->lpt_cap = 0;
        HIP_OK(hipMalloc(&c->lpt_order, count * sizeof(long long)));
        HIP_OK(hipMalloc(&c->lpt_bucket, count));
        c->lpt_cap = count;
    }
    if (!c->lpt_hist) HIP_OK(hipMalloc(&c->lpt_hist, 2 * tg::LPT_NB * sizeof(unsigned long long)));
    int rc = lpt_order_into(c, stride, count, count_dev, list, in, c->lpt_order, c->lpt_bucket, c->lpt_hist);
    if (rc) return rc;
    *order = c->lpt_order;
    return TORCHGPU_OK;
}


constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
unsigned char* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on device `dev`, which must be current;
// taken once per device.
int noalloc_setup(torchgpu_cooling* c, int dev) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        if (c) snprintf(c->msg, sizeof c->msg, "device %d beyond the %d devices of the fixed scratch", dev,
                        NOALLOC_MAX_DEVICES);
        return TORCHGPU_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_ptr[dev]) return TORCHGPU_OK;
    void* ptr = nullptr;
    HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
        if (c) snprintf(c->msg, sizeof c->msg, "hipGetSymbolAddress gave no address for the fixed scratch");
        return TORCHGPU_ERR_HIP;
    }
    noalloc_ptr[dev] = static_cast<unsigned char*>(ptr);
    return TORCHGPU_OK;
}

int ensure_queues(torchgpu_cooling* c, long long n) {
    if (n <= c->cap) return TORCHGPU_OK;
    (void)hipStreamSynchronize(c->stream);
    if (c->mlp_list) (void)hipFree(c->mlp_list);
    if (c->fb_list) (void)hipFree(c->fb_list);
    if (c->long_list) (void)hipFree(c->long_list);
    if (c->feat) (void)hipFree(c->feat);
    if (c->y) (void)hipFree(c->y);
    if (c->sdata) (void)hipFree(c->sdata);
    c->mlp_list = c->fb_list = c->long_list = nullptr;
    c->feat = c->y = nullptr;
    c->sdata = nullptr;
    c->cap = 0;
    HIP_OK(hipMalloc(&c->mlp_list, n * sizeof(long long)));
    HIP_OK(hipMalloc(&c->fb_list, n * sizeof(long long)));
    HIP_OK(hipMalloc(&c->long_list, n * sizeof(long long)));
    HIP_OK(hipMalloc(&c->feat, n * 8 * sizeof(float)));
    HIP_OK(hipMalloc(&c->y, n * 3 * sizeof(float)));
    HIP_OK(hipMalloc(&c->sdata, n * sizeof(tg::SurrCell)));
    c->cap = n;
    return TORCHGPU_OK;
}

// Reports (and clears) the error word the cooling kernels raise.
int check_err(torchgpu_cooling* c) {
    unsigned int e = 0;
    HIP_OK(hipMemcpyAsync(&e, c->err, sizeof e, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (e) {
        static const unsigned int zero = 0;
        HIP_OK(hipMemcpyAsync(c->err, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    if (e & tg::ERR_THERMO_NSTEPS) {
        snprintf(c->msg, sizeof c->msg, "Thermodynamics::integrate: subcycle count out of int range");
        return TORCHGPU_ERR_NSTEPS_RANGE;
    }
    return TORCHGPU_OK;
}

}  // namespace

extern "C" {

int torchgpu_cooling_create(const torchgpu_thermo_params* p, void* stream, torchgpu_cooling** out) {
    if (!p || !out) return TORCHGPU_ERR_INVALID_ARG;
    *out = nullptr;
    torchgpu_cooling* c = new torchgpu_cooling();
    c->p = *p;
    c->t = to_consts(*p);
    c->stream = (hipStream_t)stream;  // NULL: the null (default) stream
    double tab[78];
    for (int i = 0; i < 26; ++i) { tab[i] = p->spl_x[i]; tab[26 + i] = p->spl_y[i]; tab[52 + i] = p->spl_y2[i]; }
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        c->nsm = prop.multiProcessorCount;
    c->device = dev;
    // the fixed scratch of torchgpu_thermo_cells_noalloc on this device
    if (noalloc_setup(c, dev) != TORCHGPU_OK) {
        torchgpu_cooling_destroy(c);
        return TORCHGPU_ERR_HIP;
    }
    unsigned int zero = 0;
    if (hipMalloc(&c->tables, sizeof tab) != hipSuccess || hipMalloc(&c->counters, 6 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&c->err, sizeof(unsigned int)) != hipSuccess ||
        hipMemcpy(c->tables, tab, sizeof tab, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->err, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) {
        torchgpu_cooling_destroy(c);
        return TORCHGPU_ERR_HIP;
    }
    *out = c;
    return TORCHGPU_OK;
}

void torchgpu_cooling_destroy(torchgpu_cooling* c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    for (void* ptr : {(void*)c->tables, (void*)c->counters, (void*)c->err, (void*)c->mlp_list, (void*)c->fb_list, (void*)c->long_list,
                      (void*)c->feat, (void*)c->y, (void*)c->sdata, (void*)c->in9, (void*)c->out5, (void*)c->gpath, (void*)c->rcache,
                      (void*)c->dtbits,
                      (void*)c->lpt_order, (void*)c->lpt_bucket, (void*)c->lpt_hist, (void*)c->dcol,
                      (void*)c->col, (void*)c->xface, (void*)c->tmin})
        if (ptr) (void)hipFree(ptr);
    delete c;
}

int torchgpu_cooling_last_error(const torchgpu_cooling* c, char* buf, size_t len) {
    if (!c || !buf || !len) return TORCHGPU_ERR_INVALID_ARG;
    snprintf(buf, len, "%s", c->msg);
    return TORCHGPU_OK;
}

int torchgpu_thermo_cells(torchgpu_cooling* c, int64_t n, const double* in, double* out, int64_t* total_substeps) {
    if (!c || n < 0 || (n > 0 && (!in || !out))) return TORCHGPU_ERR_INVALID_ARG;
    if (n == 0) {
        if (total_substeps) *total_substeps = 0;
        return TORCHGPU_OK;
    }
    int rc = zero_counters(c);
    if (rc) return rc;
    const long long* order = nullptr;
    if ((rc = lpt_order(c, n, n, nullptr, nullptr, in, &order))) return rc;
    rc = launch_fp64(c, n, n, nullptr, order, in, out, nullptr);
    if (rc) return rc;
    if (total_substeps) {
        unsigned long long tot = 0;
        HIP_OK(hipMemcpyAsync(&tot, c->counters + 1, sizeof tot, hipMemcpyDeviceToHost, c->stream));
        rc = check_err(c);
        *total_substeps = (int64_t)tot;
        if (rc) return rc;
    }
    return TORCHGPU_OK;
}

int torchgpu_thermo_cells_noalloc(torchgpu_cooling* c, int64_t n, const double* in, double* out,
                                  int64_t* total_substeps) {
    if (!c || n < 0 || (n > 0 && (!in || !out))) return TORCHGPU_ERR_INVALID_ARG;
    if (n > NOALLOC_MAX_CELLS) {
        snprintf(c->msg, sizeof c->msg, "torchgpu_thermo_cells_noalloc: %lld cells, at most %lld", (long long)n,
                 NOALLOC_MAX_CELLS);
        return TORCHGPU_ERR_INVALID_ARG;
    }
    if (n == 0) {
        if (total_substeps) *total_substeps = 0;
        return TORCHGPU_OK;
    }
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    if (dev != c->device) {
        snprintf(c->msg, sizeof c->msg, "torchgpu_thermo_cells_noalloc: device %d is current, the handle is on %d",
                 dev, c->device);
        return TORCHGPU_ERR_INVALID_ARG;
    }
    int rc = noalloc_setup(c, dev);
    if (rc) return rc;
    unsigned char* base = noalloc_ptr[dev];
    unsigned long long* hist = reinterpret_cast<unsigned long long*>(base);
    long long* order = reinterpret_cast<long long*>(base + NOALLOC_HIST_BYTES);
    unsigned char* bucket = base + NOALLOC_HIST_BYTES + NOALLOC_ORDER_BYTES;
    if ((rc = zero_counters(c))) return rc;
    if ((rc = lpt_order_into(c, n, n, nullptr, nullptr, in, order, bucket, hist))) return rc;
    if ((rc = launch_fp64(c, n, n, nullptr, order, in, out, nullptr))) return rc;
    if (total_substeps) {
        unsigned long long tot = 0;
        HIP_OK(hipMemcpyAsync(&tot, c->counters + 1, sizeof tot, hipMemcpyDeviceToHost, c->stream));
        rc = check_err(c);
        *total_substeps = (int64_t)tot;
        if (rc) return rc;
    }
    return TORCHGPU_OK;
}

int torchgpu_thermo_subcycles(torchgpu_cooling* c, int64_t n, const double* in, int64_t* count) {
    if (!c || n < 0 || (n > 0 && (!in || !count))) return TORCHGPU_ERR_INVALID_ARG;
    if (n == 0) return TORCHGPU_OK;
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(tg::subcycle_count_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, c->t, c->tables, n, in,
                       (long long*)count);
    HIP_OK(hipGetLastError());
    return TORCHGPU_OK;
}

int torchgpu_cooling_set_surrogate(torchgpu_cooling* c, torchgpu_mlp* m, const torchgpu_surrogate_cfg* cfg) {
    if (!c || !cfg) return TORCHGPU_ERR_INVALID_ARG;
    c->mlp = m;
    for (int i = 0; i < 8; ++i) { c->sc.mu[i] = cfg->mu[i]; c->sc.isd[i] = cfg->isd[i]; }
    for (int i = 0; i < 7; ++i) { c->sc.lo[i] = cfg->lo[i]; c->sc.hi[i] = cfg->hi[i]; }
    c->sc.gate = cfg->gate;
    c->sc.nexact = cfg->nexact;
    c->sc.train = 0;
    c->has_surrogate = (m != nullptr);
    return TORCHGPU_OK;
}

}  // extern "C"

// Where a grid surrogate step writes the cooling update of the cells its
// prepass finishes (surrogate_prepass_kernel<true, true>): the state, its float Q
// (NULL for FP64 grids) and the step state.
struct ApplyDst {
    double* U;
    float* qf;
    tg::StepState* st;
    const double* pre_rate;  // heating rate and pre-step RATE of each cell on this state, or NULL
};

// The surrogate step on `in` (gs and apply NULL), or on the grid state (gs
// and apply both given: the inputs are formed in the prepass, `in` receives
// those of the queued cells, and the prepass applies the update of the cells
// it finishes; the caller applies the rest, marked in `path`).
static int surrogate_step(torchgpu_cooling* c, int64_t n, double* in, double* out, uint8_t* path, int64_t* counts,
                          const tg::GridSrc* gs, const ApplyDst* apply = nullptr) {
    if ((gs == nullptr) != (apply == nullptr)) return TORCHGPU_ERR_INVALID_ARG;
    if (!c->has_surrogate) return TORCHGPU_ERR_NOT_TRAINED;
    if (n == 0) return TORCHGPU_OK;
    int rc = ensure_queues(c, n);
    if (rc) return rc;
    if ((rc = zero_counters(c))) return rc;
    const uint16_t* w16 = nullptr;
    const float* b = nullptr;
    if ((rc = torchgpu_mlp_device_weights(c->mlp, &w16, &b))) return rc;
    tg::SurrogateConsts sc = c->sc;
    sc.train = 0;
    long long nb = (n + 255) / 256;
    if (nb > (long long)c->nsm * 16) nb = (long long)c->nsm * 16;
    if (gs)
        hipLaunchKernelGGL((tg::surrogate_prepass_kernel<true, true>), dim3((unsigned)nb), dim3(256), 0, c->stream, c->t,
                           c->tables, sc, (long long)n, in, out, path, c->mlp_list, c->feat, c->sdata, (float*)nullptr,
                           c->fb_list, c->counters, c->err, *gs, apply->U, apply->qf, apply->st, c->long_list,
                           apply->pre_rate);
    else
        hipLaunchKernelGGL((tg::surrogate_prepass_kernel<false, true>), dim3((unsigned)nb), dim3(256), 0, c->stream,
                           c->t, c->tables, sc, (long long)n, in, out, path, c->mlp_list, c->feat, c->sdata,
                           (float*)nullptr, c->fb_list, c->counters, c->err, tg::GridSrc{}, (double*)nullptr,
                           (float*)nullptr, (tg::StepState*)nullptr, c->long_list, (const double*)nullptr);
    HIP_OK(hipGetLastError());
    // the cells the prepass deferred (many remaining subcycles), densely
    long long nb64 = (n + 63) / 64;
    if (nb64 > (long long)c->nsm * 64) nb64 = (long long)c->nsm * 64;
    hipLaunchKernelGGL(tg::surrogate_long_kernel, dim3((unsigned)nb64), dim3(64), 0, c->stream, c->t, c->tables, sc,
                       (long long)n, (const double*)in, path, c->mlp_list, c->feat, c->sdata, c->fb_list, c->counters,
                       (const long long*)c->long_list, c->err);
    HIP_OK(hipGetLastError());
    HIP_OK(tg::launch_mlp_infer_dev(w16, b, c->feat, c->counters + 2, n, c->y, c->nsm, c->stream));
    hipLaunchKernelGGL(tg::surrogate_finish_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, c->t, sc,
                       (long long)n, c->counters, c->mlp_list, c->sdata, c->y, out, path, c->fb_list, c->counters);
    HIP_OK(hipGetLastError());
    // fallback queue through the FP64 integrator (work counter reset first)
    hipLaunchKernelGGL(tg::zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, c->counters, 1);
    HIP_OK(hipGetLastError());
    const long long* order = nullptr;
    if ((rc = lpt_order(c, n, n, c->counters + 3, c->fb_list, in, &order))) return rc;
    if ((rc = launch_fp64(c, n, n, c->counters + 3, order, in, out, path))) return rc;
    if (counts) {
        unsigned long long h[4];
        HIP_OK(hipMemcpyAsync(h, c->counters, sizeof h, hipMemcpyDeviceToHost, c->stream));
        if ((rc = check_err(c))) return rc;
        counts[0] = (int64_t)h[2];
        counts[1] = (int64_t)h[3];
    }
    return TORCHGPU_OK;
}

extern "C" {

int torchgpu_thermo_cells_surrogate(torchgpu_cooling* c, int64_t n, const double* in, double* out, uint8_t* path,
                                    int64_t* counts) {
    if (!c || n < 0 || (n > 0 && (!in || !out))) return TORCHGPU_ERR_INVALID_ARG;
    // the prepass only reads `in` when it is not forming the inputs itself
    return surrogate_step(c, n, const_cast<double*>(in), out, path, counts, nullptr);
}

int torchgpu_surrogate_training_data(torchgpu_cooling* c, int64_t n, const double* in, const double* out,
                                     int64_t* list, float* feat, float* target, float* raw, int64_t* count) {
    if (!c || n <= 0 || !in || !out || !list || !feat || !target || !count) return TORCHGPU_ERR_INVALID_ARG;
    int rc = ensure_queues(c, n);
    if (rc) return rc;
    if ((rc = zero_counters(c))) return rc;
    tg::SurrogateConsts sc = c->sc;
    sc.train = 1;
    // the prepass rewrites `out` for the exactly-integrated cells with the
    // same FP64 values the integrator produced; a scratch copy keeps `out`
    // untouched
    double* scratch = nullptr;
    HIP_OK(hipMalloc(&scratch, 5 * n * sizeof(double)));
    // every path below frees scratch: the first HIP error is kept and
    // reported after the free
    hipError_t e = hipSuccess;
    auto keep = [&e](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    long long nb = (n + 255) / 256;
    if (nb > (long long)c->nsm * 16) nb = (long long)c->nsm * 16;
    hipLaunchKernelGGL((tg::surrogate_prepass_kernel<false, false>), dim3((unsigned)nb), dim3(256), 0, c->stream,
                       c->t, c->tables, sc, (long long)n, const_cast<double*>(in), scratch, (unsigned char*)nullptr,
                       (long long*)list, feat, c->sdata, raw, c->fb_list, c->counters, c->err, tg::GridSrc{},
                       (double*)nullptr, (float*)nullptr, (tg::StepState*)nullptr, (long long*)nullptr,
                       (const double*)nullptr);
    keep(hipGetLastError());
    if (e == hipSuccess) {
        hipLaunchKernelGGL(tg::surrogate_targets_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, c->counters,
                           (long long)n, (const long long*)list, c->sdata, out, target);
        keep(hipGetLastError());
    }
    unsigned long long h[4] = {0, 0, 0, 0};
    if (e == hipSuccess) keep(hipMemcpyAsync(h, c->counters, sizeof h, hipMemcpyDeviceToHost, c->stream));
    keep(hipStreamSynchronize(c->stream));
    keep(hipFree(scratch));
    HIP_OK(e);
    *count = (int64_t)h[2];
    return check_err(c);
}

static int ensure_grid(torchgpu_cooling* c, long long n) {
    if (n <= c->grid_cap) return TORCHGPU_OK;
    (void)hipStreamSynchronize(c->stream);
    if (c->in9) (void)hipFree(c->in9);
    if (c->out5) (void)hipFree(c->out5);
    if (c->gpath) (void)hipFree(c->gpath);
    c->in9 = c->out5 = nullptr;
    c->gpath = nullptr;
    c->grid_cap = 0;
    HIP_OK(hipMalloc(&c->in9, 9 * n * sizeof(double)));
    HIP_OK(hipMalloc(&c->out5, 5 * n * sizeof(double)));
    HIP_OK(hipMalloc(&c->gpath, n));
    // finite contents everywhere: the ray trace reads columns of zero weight
    hipLaunchKernelGGL(tg::zero_f64_kernel, dim3(1024), dim3(256), 0, c->stream, c->in9, 9 * n);
    HIP_OK(hipGetLastError());
    if (!c->dtbits) HIP_OK(hipMalloc(&c->dtbits, sizeof(unsigned long long)));
    c->grid_cap = n;
    return TORCHGPU_OK;
}

// The grid source of the cooling inputs (see tg::GridSrc) for the
// operator's current set-up.
static int grid_src(torchgpu_cooling* c, torchgpu_ctx* ctx, const double* T_min, const double* col_den, const double* F,
                    long long excluded, double dt, tg::GridSrc& gs) {
    tg::CtxView v = tg::ctx_view(ctx);
    const long long n = v.g->ncell;
    if (c->star.on && (c->star.n[0] != v.g->nx || c->star.n[1] != v.g->ny || c->star.n[2] != v.g->nz))
        return TORCHGPU_ERR_INVALID_ARG;
    if (!T_min && c->tmin) {
        if (c->tmin_n != n) return TORCHGPU_ERR_INVALID_ARG;
        T_min = c->tmin;
    }
    int rc = ensure_grid(c, n);
    if (rc) return rc;
    if (c->star.on && !col_den) {
        // Thermodynamics::preTimeStepCalculations starts with rayTrace on
        // the fixed primitives (Thermodynamics.cpp:218-220): the density of
        // U with fixPrimitives' floor
        HIP_OK(tg::launch_thermo_ray_trace(c->star, v.U, v.g->dfloor, c->col, c->dcol, c->xface, c->stream));
        col_den = c->col;
    }
    gs = tg::GridSrc{};
    gs.g = *v.g;
    gs.U = v.U;
    gs.T_min = T_min;
    gs.T_min_default = 100.0;
    gs.col_den = col_den;
    gs.F = F;
    gs.dt = dt;
    gs.st = v.st;
    gs.excluded = excluded;
    gs.star = c->star;
    return TORCHGPU_OK;
}

static int gather(torchgpu_cooling* c, torchgpu_ctx* ctx, const double* T_min, const double* col_den, const double* F,
                  long long excluded, double dt) {
    tg::GridSrc gs;
    int rc = grid_src(c, ctx, T_min, col_den, F, excluded, dt, gs);
    if (rc) return rc;
    const long long n = gs.g.ncell;
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(tg::thermo_gather_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, gs, c->in9);
    HIP_OK(hipGetLastError());
    return TORCHGPU_OK;
}

// Thermodynamics::calculateTimeStep on the grid into c->dtbits (bit pattern
// of min |f E/RATE| over the gated cells, 0x7fef... when none), no host sync.
// The thermo time step of the grid state; with cfl_st, the kernel also
// reduces the hydro CFL maximum of the same state (launch_cfl's); with
// keep_rate, it keeps each cell's heating rate and pre-step RATE in
// c->rcache for a cooling sub-step that follows on the same state.
static int grid_timestep_dev(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                             const double* F, int64_t excluded, tg::StepState* cfl_st = nullptr,
                             bool keep_rate = false) {
    // the inputs are formed in the kernel
    tg::GridSrc gs;
    int rc = grid_src(c, ctx, T_min, col_den, F, excluded, 0.0, gs);
    if (rc) return rc;
    const long long n = tg::ctx_view(ctx).g->ncell;
    if (keep_rate && n > c->rcache_cap) {
        (void)hipStreamSynchronize(c->stream);
        if (c->rcache) (void)hipFree(c->rcache);
        c->rcache = nullptr;
        c->rcache_cap = 0;
        HIP_OK(hipMalloc(&c->rcache, 2 * n * sizeof(double)));
        c->rcache_cap = n;
    }
    hipLaunchKernelGGL(tg::set_u64_kernel, dim3(1), dim3(1), 0, c->stream, c->dtbits, 0x7fefffffffffffffull);
    long long nb = (n + 255) / 256;
    if (nb > (long long)c->nsm * 16) nb = (long long)c->nsm * 16;
    hipLaunchKernelGGL(tg::thermo_timestep_kernel<true>, dim3((unsigned)nb), dim3(256), 0, c->stream, c->t,
                       c->tables, n, (const double*)nullptr, c->dtbits, c->err, gs, cfl_st,
                       keep_rate ? c->rcache : (double*)nullptr);
    HIP_OK(hipGetLastError());
    return TORCHGPU_OK;
}

// One thermo sub-step on the grid (dt explicit or a DT_* code), no host sync.
// kept_rate: the state is the one grid_timestep_dev(keep_rate) last saw,
// so the surrogate prepass takes its heating rates and pre-step RATEs.
static int grid_substep_dev(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                            const double* F, int64_t excluded, double dt, int surrogate, bool kept_rate = false) {
    tg::CtxView v = tg::ctx_view(ctx);
    const long long n = v.g->ncell;
    // the surrogate step forms the inputs in its prepass unless the heating
    // record needs the gathered array
    const bool direct = surrogate && !(v.heat && c->t.subcycling);
    int rc;
    if (direct) {
        // the inputs formed in the prepass, which also applies the update
        // (Thermodynamics::updateSourceTerms + advSolution(dt) +
        // fixSolution) to the cells it finishes; the queued cells get it
        // once the network and the fallback have run
        tg::GridSrc gs;
        if ((rc = grid_src(c, ctx, T_min, col_den, F, excluded, dt, gs))) return rc;
        float* qf = tg::ctx_qf_for_state(ctx);
        const ApplyDst ad{v.U, qf, v.st, kept_rate ? c->rcache : nullptr};
        if ((rc = surrogate_step(c, n, c->in9, c->out5, c->gpath, nullptr, &gs, &ad))) return rc;
        long long nb = (n + 255) / 256;
        if (nb > 65536) nb = 65536;
        hipLaunchKernelGGL(tg::queue_apply_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, gs, c->out5, c->gpath,
                           v.U, qf, v.st, c->err, (const unsigned long long*)c->counters);
        HIP_OK(hipGetLastError());
        if (qf) tg::ctx_qf_written(ctx);
        else tg::ctx_state_changed(ctx);
        return TORCHGPU_OK;
    } else {
        if ((rc = gather(c, ctx, T_min, col_den, F, excluded, dt))) return rc;
        if (surrogate) rc = surrogate_step(c, n, c->in9, c->out5, nullptr, nullptr, nullptr);
        else rc = torchgpu_thermo_cells(c, n, c->in9, c->out5, nullptr);
    }
    if (rc) return rc;
    if (v.heat && c->t.subcycling) {
        // GridCell::H[TOT] of Thermodynamics::integrate (zeroed below the switch)
        long long nb = (n + 255) / 256;
        if (nb > 65536) nb = 65536;
        hipLaunchKernelGGL(tg::heat_integrate_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, n, c->in9, c->out5,
                           dt, v.st, c->t.thermo_hii_switch, v.heat);
        HIP_OK(hipGetLastError());
    }
    // Thermodynamics::updateSourceTerms + advSolution(dt) + fixSolution; a
    // mixed-precision grid also gets the float Q of the result here
    float* qf = tg::ctx_qf_for_state(ctx);
    if (qf) {
        HIP_OK(tg::launch_fix_solution_qf(*v.g, v.U, c->out5 + 2 * n, dt, v.st, c->stream, c->err, qf));
        tg::ctx_qf_written(ctx);
    } else {
        HIP_OK(tg::launch_fix_solution(*v.g, v.U, c->out5 + 2 * n, dt, v.st, c->stream, c->err));
        tg::ctx_state_changed(ctx);
    }
    return TORCHGPU_OK;
}

int torchgpu_thermo_grid_timestep(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                                  const double* F, int64_t excluded, double dt_max, double* dt) {
    if (!ctx || !c || !dt) return TORCHGPU_ERR_INVALID_ARG;
    int rc = grid_timestep_dev(ctx, c, T_min, col_den, F, excluded);
    if (rc) return rc;
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, c->dtbits, sizeof bits, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    double m;
    std::memcpy(&m, &bits, sizeof m);
    // Thermodynamics.cpp:465-474: dt = dt_max, lowered by every smaller dti
    *dt = (m < dt_max) ? m : dt_max;
    return TORCHGPU_OK;
}

int torchgpu_thermo_grid_substep(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                                 const double* F, int64_t excluded, double dt, int surrogate) {
    if (!ctx || !c || !(dt >= 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    if (surrogate && !c->has_surrogate) return TORCHGPU_ERR_NOT_TRAINED;
    return grid_substep_dev(ctx, c, T_min, col_den, F, excluded, dt, surrogate);
}

int torchgpu_fill_heating(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* col_den, const double* F,
                          int64_t excluded, double* H, int on_device) {
    if (!ctx || !c || !H) return TORCHGPU_ERR_INVALID_ARG;
    tg::CtxView v = tg::ctx_view(ctx);
    if (!v.heat) {
        snprintf(c->msg, sizeof c->msg, "fill_heating: the grid does not track its heating terms "
                                        "(torchgpu_heating_track before the run)");
        return TORCHGPU_ERR_INVALID_ARG;
    }
    const long long n = v.g->ncell;
    if (c->star.on && (c->star.n[0] != v.g->nx || c->star.n[1] != v.g->ny || c->star.n[2] != v.g->nz))
        return TORCHGPU_ERR_INVALID_ARG;
    const double* Q = tg::ctx_print_state_dev(ctx);
    if (!Q) {
        snprintf(c->msg, sizeof c->msg, "fill_heating: the printed state is not available");
        return TORCHGPU_ERR_INVALID_ARG;
    }
    int rc = ensure_grid(c, n);
    if (rc) return rc;
    if (c->star.on && !col_den) {
        // fillHeatingArrays starts with rayTrace on Fluid::Q (Thermodynamics.cpp:420-421)
        HIP_OK(tg::launch_thermo_ray_trace(c->star, Q, -HUGE_VAL, c->col, c->dcol, c->xface, c->stream));
        col_den = c->col;
    }
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    if (on_device) {
        hipLaunchKernelGGL(tg::fill_heating_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, *v.g, c->t, c->tables,
                           Q, col_den, F, c->star, (long long)excluded, v.heat, H);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(c->stream));
        return TORCHGPU_OK;
    }
    // host output: the 11 planes through a temporary device array, freed on
    // every path
    double* tmp = nullptr;
    HIP_OK(hipMalloc(&tmp, 11 * n * sizeof(double)));
    hipError_t e = hipSuccess;
    hipLaunchKernelGGL(tg::fill_heating_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, *v.g, c->t, c->tables, Q,
                       col_den, F, c->star, (long long)excluded, v.heat, tmp);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(H, tmp, 11 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = es;
    const hipError_t ef = hipFree(tmp);
    if (e == hipSuccess) e = ef;
    HIP_OK(e);
    return TORCHGPU_OK;
}

int torchgpu_cooling_set_star(torchgpu_cooling* c, torchgpu_ctx* ctx, const torchgpu_star* star) {
    if (!c || !ctx) return TORCHGPU_ERR_INVALID_ARG;
    if (!star) {
        c->star = tg::StarGeom{};
        return TORCHGPU_OK;
    }
    const tg::HydroGeom& g = *tg::ctx_view(ctx).g;
    const int n[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a)
        if (star->cell[a] < 0 || star->cell[a] >= n[a]) {
            snprintf(c->msg, sizeof c->msg, "star cell (%d, %d, %d) outside the %d x %d x %d grid", star->cell[0],
                     star->cell[1], star->cell[2], n[0], n[1], n[2]);
            return TORCHGPU_ERR_INVALID_ARG;
        }
    if (star->wind_cell_radius < 0 || !(star->photon_rate >= 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    if (g.ncell >= (1LL << 32)) {  // 32-bit index math of the shell sweeps
        snprintf(c->msg, sizeof c->msg, "star operators need fewer than 2^32 cells");
        return TORCHGPU_ERR_INVALID_ARG;
    }
    const long long ncell = g.ncell;
    if (ncell > c->dcol_cap) {
        (void)hipStreamSynchronize(c->stream);
        if (c->dcol) (void)hipFree(c->dcol);
        if (c->col) (void)hipFree(c->col);
        c->dcol = c->col = nullptr;
        c->dcol_cap = 0;
        HIP_OK(hipMalloc(&c->dcol, ncell * sizeof(double)));
        HIP_OK(hipMalloc(&c->col, ncell * sizeof(double)));
        c->dcol_cap = ncell;
    }
    hipLaunchKernelGGL(tg::zero_f64_kernel, dim3(1024), dim3(256), 0, c->stream, c->dcol, ncell);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(tg::zero_f64_kernel, dim3(1024), dim3(256), 0, c->stream, c->col, ncell);
    HIP_OK(hipGetLastError());
    tg::StarGeom s{};
    s.on = 1;
    for (int a = 0; a < 3; ++a) {
        s.s[a] = star->cell[a];
        s.n[a] = n[a];
    }
    s.wind_r = star->wind_cell_radius;
    s.dx = g.dx;
    s.photon_rate = star->photon_rate;
    s.pi = c->p.pi;
    s.hydrogen_mass = c->p.hydrogen_mass;
    const long long nx = tg::xface_doubles(s);
    if (nx > c->xface_cap) {
        (void)hipStreamSynchronize(c->stream);
        if (c->xface) (void)hipFree(c->xface);
        c->xface = nullptr;
        c->xface_cap = 0;
        HIP_OK(hipMalloc(&c->xface, nx * sizeof(double)));
        c->xface_cap = nx;
    }
    c->star = s;
    return TORCHGPU_OK;
}

int torchgpu_cooling_set_min_temp(torchgpu_cooling* c, const double* T_min, int64_t n, int on_device) {
    if (!c || (T_min && n < 1)) return TORCHGPU_ERR_INVALID_ARG;
    HIP_OK(hipStreamSynchronize(c->stream));
    if (c->tmin) (void)hipFree(c->tmin);
    c->tmin = nullptr;
    c->tmin_n = 0;
    if (!T_min) return TORCHGPU_OK;
    HIP_OK(hipMalloc(&c->tmin, n * sizeof(double)));
    c->tmin_n = n;
    HIP_OK(hipMemcpyAsync(c->tmin, T_min, n * sizeof(double),
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return TORCHGPU_OK;
}

int torchgpu_star_geometry(torchgpu_cooling* c, int64_t* nb, double* w, double* ds, double* shell_vol, int on_device) {
    if (!c || !nb || !w || !ds || !shell_vol || !c->star.on) return TORCHGPU_ERR_INVALID_ARG;
    const long long n = (long long)c->star.n[0] * c->star.n[1] * c->star.n[2];
    long long* d_nb = reinterpret_cast<long long*>(nb);
    double *d_w = w, *d_ds = ds, *d_sv = shell_vol;
    void* tmp = nullptr;
    if (!on_device) {
        HIP_OK(hipMalloc(&tmp, n * (4 * sizeof(long long) + 6 * sizeof(double))));
        d_nb = static_cast<long long*>(tmp);
        d_w = reinterpret_cast<double*>(d_nb + 4 * n);
        d_ds = d_w + 4 * n;
        d_sv = d_ds + n;
    }
    hipError_t e = tg::launch_star_geometry(c->star, d_nb, d_w, d_ds, d_sv, c->stream);
    if (e == hipSuccess && !on_device) {
        e = hipMemcpyAsync(nb, d_nb, 4 * n * sizeof(long long), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(w, d_w, 4 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(ds, d_ds, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(shell_vol, d_sv, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tmp) (void)hipFree(tmp);
    HIP_OK(e);
    return TORCHGPU_OK;
}

int torchgpu_thermo_grid_inputs(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                                const double* F, int64_t excluded, double* in9, double* dcol_den, int on_device) {
    if (!ctx || !c || !in9) return TORCHGPU_ERR_INVALID_ARG;
    int rc = gather(c, ctx, T_min, col_den, F, excluded, 0.0);
    if (rc) return rc;
    const long long n = tg::ctx_view(ctx).g->ncell;
    const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_OK(hipMemcpyAsync(in9, c->in9, 9 * n * sizeof(double), k, c->stream));
    if (dcol_den && c->star.on && !col_den)
        HIP_OK(hipMemcpyAsync(dcol_den, c->dcol, n * sizeof(double), k, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return check_err(c);
}

int torchgpu_run_split(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                       const double* F, int64_t excluded, int nsteps, int surrogate) {
    if (!c) return TORCHGPU_ERR_INVALID_ARG;
    return torchgpu_run_full(ctx, c, nullptr, T_min, col_den, F, excluded, nsteps, surrogate);
}

// The component of a step's first sub-step: comps rotated by `launched`.
static int order_first(const int* comps, int ncomps, int launched) { return comps[launched % ncomps]; }

int torchgpu_run_full(torchgpu_ctx* ctx, torchgpu_cooling* c, torchgpu_radiation* r, const double* T_min,
                      const double* col_den, const double* F, int64_t excluded, int nsteps, int surrogate) {
    if (!ctx || nsteps < 0) return TORCHGPU_ERR_INVALID_ARG;
    if (!c && !r) return torchgpu_run(ctx, nsteps);  // one component: Torch::hydroStep
    if (c && surrogate && !c->has_surrogate) return TORCHGPU_ERR_NOT_TRAINED;
    tg::CtxView v = tg::ctx_view(ctx);
    // one stream orders the handles' work
    if ((c && v.stream != c->stream) || (r && v.stream != tg::rad_stream(r))) return TORCHGPU_ERR_INVALID_ARG;
    tg::ctx_set_print_mode(ctx, r ? 3 : 2);
    double* print_state = tg::ctx_print_buffer(ctx);
    // activeComponents (Torch.cpp:256-261): hydro, thermodynamics, radiation
    enum { HYDRO, THERMO, RAD };
    int comps[3], ncomps = 0;
    comps[ncomps++] = HYDRO;
    if (c) comps[ncomps++] = THERMO;
    if (r) comps[ncomps++] = RAD;
    // Torch::fullStep's sub-steps (Torch.cpp:397-415): h dt (h = 1/2, ...,
    // 1 for the last) over the components rotated by stepCounter, then dt/2
    // back down; the first reuses the pre-step calculations
    int order[5];
    double code[5];
    const int nsub = 2 * ncomps - 1;
    int rc;
    // Torch::stepCounter turns once per step that ran: bring it up to date
    // with the steps the device completed since the last call (the host
    // launched more when a pause or tmax halted the run)
    tg::StepState st_now;
    HIP_OK(hipMemcpyAsync(&st_now, v.st, sizeof st_now, hipMemcpyDeviceToHost, v.stream));
    HIP_OK(hipStreamSynchronize(v.stream));
    int& counter = tg::ctx_step_counter(ctx);
    int& counter_step = tg::ctx_counter_step(ctx);
    counter = (counter + (st_now.step - counter_step)) % ncomps;
    counter_step = st_now.step;
    int launched = counter;
    for (int s = 0; s < nsteps; ++s) {
        v = tg::ctx_view(ctx);  // U and U* swap roles after every hydro sub-step
        // Torch::fullStep's preTimeStepCalculations and calculateTimeStep
        // (Torch.cpp:304-339,388-397) on the device: the hydro CFL maximum,
        // the thermo and radiation minima, then dt in step_begin; with the
        // thermo component the CFL maximum comes from the thermo time step
        // kernel, which reads the same state
        if (!c) HIP_OK(tg::launch_cfl(*v.g, v.U, v.st, v.stream));
        // a step that opens with the cooling sub-step (every other one: the
        // order turns each step) reuses the pre-step rates of the time step,
        // which sees the same state (not with radiation, whose pre-step
        // comes in between)
        const bool share = c && !r && surrogate && order_first(comps, ncomps, (launched + 1) % ncomps) == THERMO;
        if (c && (rc = grid_timestep_dev(ctx, c, T_min, col_den, F, excluded, v.st, share))) return rc;
        if (r && ((rc = tg::rad_pre_step_dev(r, 1)) || (rc = tg::rad_timestep_dev(r, 0.0, 1)))) return rc;
        HIP_OK(tg::launch_step_begin(v.st, 1, 0.0, v.stream, c ? c->dtbits : nullptr,
                                     r ? tg::rad_dt_bits(r) : nullptr));
        if (r && (rc = tg::rad_once_commit(r, 1))) return rc;
        launched = (launched + 1) % ncomps;
        for (int i = 0; i < ncomps; ++i) {
            order[i] = comps[(i + launched) % ncomps];
            code[i] = (i == ncomps - 1) ? tg::DT_FULL : tg::DT_HALF;
        }
        for (int i = ncomps - 2, j = ncomps; i >= 0; --i, ++j) {
            order[j] = comps[(i + launched) % ncomps];
            code[j] = tg::DT_HALF;
        }
        for (int k = 0; k < nsub; ++k) {
            const bool last = k == nsub - 1;
            if (last && print_state) {
                // the reference's Fluid::Q after the step is fixPrimitives(Q)
                // of the last sub-step's input (Torch.cpp:352-358), kept for
                // the snapshots printed at checkpoints
                tg::CtxView w = tg::ctx_view(ctx);
                if (r) HIP_OK(tg::launch_q_from_u_if_running(*w.g, w.st, w.U, print_state, w.stream));
                else HIP_OK(tg::launch_copy_if_running(w.st, w.U, print_state, 7 * w.g->ncell, w.stream));
            }
            if (order[k] == HYDRO) {
                if ((rc = tg::ctx_run_stage(ctx, 0, code[k]))) return rc;
            } else if (order[k] == THERMO) {
                if ((rc = grid_substep_dev(ctx, c, T_min, col_den, F, excluded, code[k], surrogate, share && k == 0)))
                    return rc;
            } else {
                if ((rc = tg::rad_substep_dev(r, code[k], k == 0))) return rc;
                // Radiation::integrate leaves its HII fraction in Fluid::Q
                if (last && print_state && (rc = tg::rad_print_hii(r, print_state))) return rc;
            }
        }
        HIP_OK(tg::launch_step_end(tg::ctx_view(ctx).st, v.stream));
    }
    return TORCHGPU_OK;
}

}  // extern "C"
