// api.hip -- C ABI over the HIP kernels (see include/torchgpu.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>

#include "../../include/torchgpu.h"
#include "torchgpu_internal.hpp"

struct torchgpu_ctx {
    torchgpu_params p;
    tg::HydroGeom g;
    hipStream_t stream = nullptr;
    double* U = nullptr;      // conserved state
    double* Ustar = nullptr;  // predictor state (also scratch for Q transfers)
    tg::StepState* st = nullptr;
    // mixed path: float primitives, qf[qcur] = fixPrimitives(QfromU(U)) with
    // the "hydro before" marks when qvalid, qf[1 - qcur] = QfromU(U*) when
    // qstar_valid; the stage kernels write the next stage's
    float* qf[2] = {nullptr, nullptr};
    int qcur = 0;
    bool qvalid = false, qstar_valid = false;
    int step_counter = 0;  // Torch::stepCounter of the operator-split loop
    int counter_step = 0;  // ... as of this device step count
    // the primitive state the reference's Fluid holds after the last step,
    // which Torch::run prints: 1 = QfromU(U*) (torchgpu_run, hydro_step),
    // 2 = fixPrimitives(QfromU(U*)) with U* the input of the last sub-step
    // (torchgpu_run_split); before any step, fixPrimitives(QfromU(U))
    int last_mode = 0;
    int state_step = 0;   // StepState::step when U was last set from outside
    bool pause_ckpt = false;
    double* qprint = nullptr;
    double* pstate = nullptr;  // operator-split: input of the last sub-step
    double* heat = nullptr;    // GridCell::H terms kept across steps while tracked ([3][n])
    std::string err;
};

namespace {

int fail(torchgpu_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(torchgpu_ctx* c, hipError_t e, const char* where) {
    return fail(c, TORCHGPU_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, call)                                   \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

size_t state_bytes(const torchgpu_ctx* c) { return sizeof(double) * 7 * (size_t)c->g.ncell; }

// Maps the device error bits of the first failing step to the exception the
// reference throws first inside that step.
int check_device_errors(torchgpu_ctx* c) {
    tg::StepState st;
    HIP_TRY(c, hipMemcpyAsync(&st, c->st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (!st.err) return TORCHGPU_OK;
    const std::string at = " (step " + std::to_string(st.err_step1 - 1) + ")";
    const unsigned int first = st.err & (0u - st.err);
    switch (first) {
    case tg::ERR_CHECK_VALUES:
        return fail(c, TORCHGPU_ERR_CHECK_VALUES, "hydro before produced an error." + at);
    case tg::ERR_RIEMANN_NAN_1:
    case tg::ERR_RIEMANN_NAN_2:
        return fail(c, TORCHGPU_ERR_RIEMANN_NAN, "RiemannSolver::solve: HII flux is NaN" + at);
    case tg::ERR_THERMO_NSTEPS:
        return fail(c, TORCHGPU_ERR_NSTEPS_RANGE, "Thermodynamics::integrate: subcycle count out of int range" + at);
    case tg::ERR_RAD_CHECK:
        return fail(c, TORCHGPU_ERR_RADIATION, "Radiationbefore produced an error." + at);
    case tg::ERR_RAD_NAN:
        return fail(c, TORCHGPU_ERR_RADIATION, "heating rate in radiation module is NaN" + at);
    case tg::ERR_RAD_ITER:
        return fail(c, TORCHGPU_ERR_RADIATION, "Radiation::calculate_HIIfrac: implicit method not converging." + at);
    case tg::ERR_RAD_DT:
        return fail(c, TORCHGPU_ERR_RADIATION, "Radiation::calculateTimeStep(): invalid timestep" + at);
    default:
        return fail(c, TORCHGPU_ERR_FIX_SOLUTION, "Fluid::fixSolution: invalid value." + at);
    }
}

// Mixed path: the power-of-two state scale from the current state (the
// predictor kernels keep it up to date from then on).  Synchronises; also
// records that U was set from outside (for the printed state).
int seed_scale(torchgpu_ctx* c) {
    if (c->p.precision == TORCHGPU_MIXED) HIP_TRY(c, tg::launch_mixed_scale_seed(c->g, c->U, c->st, c->stream));
    tg::StepState st;
    HIP_TRY(c, hipMemcpyAsync(&st, c->st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->state_step = st.step;
    return TORCHGPU_OK;
}

int reset_clock(torchgpu_ctx* c) {
    tg::StepState st;
    std::memset(&st, 0, sizeof st);
    st.dt_max = c->p.dt_max;
    st.tmax = c->p.tmax;
    st.first = 1;
    st.debug = c->p.debug ? 1 : 0;
    st.ncheckpoints = c->p.ncheckpoints > 0 ? c->p.ncheckpoints : 1;
    c->step_counter = 0;
    c->counter_step = 0;
    HIP_TRY(c, hipMemcpyAsync(c->st, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return TORCHGPU_OK;
}

}  // namespace

namespace {

// Fills c->qprint with Fluid::Q as Torch::run prints it (see
// torchgpu_get_print_state).
int print_state(torchgpu_ctx* c) {
    tg::StepState st;
    HIP_TRY(c, hipMemcpyAsync(&st, c->st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const int mode = st.step > c->state_step ? c->last_mode : 0;
    if (!c->qprint) HIP_TRY(c, hipMalloc(&c->qprint, state_bytes(c)));
    if (mode >= 2 && !c->pstate) return fail(c, TORCHGPU_ERR_INVALID_ARG, "print state: pause at checkpoints first");
    if (mode == 0) HIP_TRY(c, tg::launch_q_from_u(c->g, c->U, c->qprint, 1, c->stream));
    else if (mode == 1) HIP_TRY(c, tg::launch_q_from_u(c->g, c->Ustar, c->qprint, 0, c->stream));
    else if (mode == 2) HIP_TRY(c, tg::launch_q_from_u(c->g, c->pstate, c->qprint, 1, c->stream));
    else HIP_TRY(c, hipMemcpyAsync(c->qprint, c->pstate, state_bytes(c), hipMemcpyDeviceToDevice, c->stream));
    return TORCHGPU_OK;
}

}  // namespace

const double* tg::ctx_print_state_dev(torchgpu_ctx* c) { return print_state(c) == TORCHGPU_OK ? c->qprint : nullptr; }

tg::CtxView tg::ctx_view(torchgpu_ctx* c) { return tg::CtxView{c->U, c->Ustar, c->st, c->stream, &c->g, c->heat}; }

void tg::ctx_state_changed(torchgpu_ctx* c) { c->qvalid = false; }
float* tg::ctx_qf_for_state(torchgpu_ctx* c) {
    return c->p.precision == TORCHGPU_FP64 ? nullptr : c->qf[c->qcur];
}
void tg::ctx_qf_written(torchgpu_ctx* c) { c->qvalid = true; }

namespace {

// One hydro stage on the context's buffers.  mode 1: predictor U* from U;
// 2: corrector U from U* and U; 0: single-stage sub-step, U* from U, then
// the two buffers swap roles.  dtf < 0: dt from the step state.
// last: the predictor of this step is the last of the caller's request, so
// its U* is stored whether or not the device loop foresees a print after it.
int run_stage(torchgpu_ctx* c, int mode, double dtf, bool last = true) {
    const int solver = c->p.riemann_solver;
    if (c->p.precision == TORCHGPU_FP64) {
        if (mode == 2) HIP_TRY(c, tg::launch_hydro_stage(c->g, solver, 2, c->Ustar, c->U, c->U, c->st, c->stream, dtf));
        else HIP_TRY(c, tg::launch_hydro_stage(c->g, solver, 1, c->U, c->U, c->Ustar, c->st, c->stream, dtf));
        if (mode == 0 && (dtf == tg::DT_HALF || dtf == tg::DT_FULL))
            HIP_TRY(c, tg::launch_halted_copy(c->g, c->st, c->U, c->Ustar, nullptr, nullptr, c->stream));
        if (mode == 0) std::swap(c->U, c->Ustar);
        return TORCHGPU_OK;
    }
    float* q0 = c->qf[c->qcur];
    float* q1 = c->qf[1 - c->qcur];
    if (mode == 2) {
        if (!c->qstar_valid) HIP_TRY(c, tg::launch_q_float(c->g, c->Ustar, q1, 0, c->stream));
        HIP_TRY(c, tg::launch_hydro_stage_mixed(c->g, solver, 2, q1, c->U, c->U, q0, 1, c->st, c->stream, dtf));
        c->qvalid = true;
        c->qstar_valid = true;
        return TORCHGPU_OK;
    }
    if (!c->qvalid) HIP_TRY(c, tg::launch_q_float(c->g, c->U, q0, 1, c->stream));
    HIP_TRY(c, tg::launch_hydro_stage_mixed(c->g, solver, 1, q0, c->U, c->Ustar, q1, mode == 0, c->st, c->stream,
                                            dtf, (mode == 0 || last) ? 1 : 0));
    if (mode == 0 && (dtf == tg::DT_HALF || dtf == tg::DT_FULL))
        HIP_TRY(c, tg::launch_halted_copy(c->g, c->st, c->U, c->Ustar, q0, q1, c->stream));
    if (mode == 1) {
        c->qvalid = true;
        c->qstar_valid = true;
    } else {
        std::swap(c->U, c->Ustar);
        c->qcur = 1 - c->qcur;
        c->qvalid = true;
        c->qstar_valid = false;
    }
    return TORCHGPU_OK;
}

}  // namespace

int tg::ctx_run_stage(torchgpu_ctx* c, int mode, double dtf) { return run_stage(c, mode, dtf); }
int& tg::ctx_step_counter(torchgpu_ctx* c) { return c->step_counter; }
int& tg::ctx_counter_step(torchgpu_ctx* c) { return c->counter_step; }
void tg::ctx_set_print_mode(torchgpu_ctx* c, int mode) { c->last_mode = mode; }
double* tg::ctx_print_buffer(torchgpu_ctx* c) {
    if (!c->pause_ckpt) return nullptr;
    if (!c->pstate && hipMalloc(&c->pstate, state_bytes(c)) != hipSuccess) return nullptr;
    return c->pstate;
}

extern "C" {

int torchgpu_create(const torchgpu_params* p, void* stream, torchgpu_ctx** out) {
    if (!p || !out) return TORCHGPU_ERR_INVALID_ARG;
    *out = nullptr;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return TORCHGPU_ERR_INVALID_ARG;
    if (p->spatial_order != 0 && p->spatial_order != 1) return TORCHGPU_ERR_INVALID_ORDER;
    if (p->riemann_solver < 0 || p->riemann_solver > 2) return TORCHGPU_ERR_UNKNOWN_SOLVER;
    if (p->slope_limiter < 0 || p->slope_limiter > 6) return TORCHGPU_ERR_UNKNOWN_LIMITER;
    for (int d = 0; d < 3; ++d)
        if (p->bc_left[d] < 0 || p->bc_left[d] > 4 || p->bc_right[d] < 0 || p->bc_right[d] > 4)
            return TORCHGPU_ERR_INVALID_ARG;
    if (p->precision != TORCHGPU_FP64 && p->precision != TORCHGPU_MIXED) return TORCHGPU_ERR_INVALID_ARG;

    torchgpu_ctx* c = new torchgpu_ctx();
    c->p = *p;
    tg::HydroGeom& g = c->g;
    std::memset(&g, 0, sizeof g);
    g.nx = p->nx; g.ny = p->ny; g.nz = p->nz;
    g.ntx = (p->nx + 7) / 8; g.nty = (p->ny + 7) / 8; g.ntz = (p->nz + 7) / 8;
    g.ntiles = g.ntx * g.nty * g.ntz;
    for (int d = 0; d < 3; ++d) { g.bc_l[d] = p->bc_left[d]; g.bc_r[d] = p->bc_right[d]; }
    g.limiter = p->slope_limiter;
    g.spatial_order = p->spatial_order;
    // TORCHGPU_ADDR64=1: the mixed kernel's 64-bit addressing (the path of
    // grids of 2^29 cells and more) on any grid, so that tests reach it
    const char* a64 = std::getenv("TORCHGPU_ADDR64");
    g.addr64 = (a64 && a64[0] == '1') ? 1 : 0;
    g.ncell = (long long)p->nx * p->ny * p->nz;
    // Grid::initialise: dx[i] = sideLength / ncells[0] (Grid.cpp:471-472)
    g.dx = p->side_length / (double)p->nx;
    // Cartesian area and volume (Grid.cpp:846-907), A/V as Hydro.cpp:129-130
    const double area = 1.0 * g.dx * g.dx;
    const double vol = 1.0 * g.dx * g.dx * g.dx;
    g.kflux = area / vol;
    g.dfloor = p->dfloor; g.pfloor = p->pfloor; g.tfloor = p->tfloor;
    g.X = p->mass_fraction_h;
    g.mu_c = (1.0 - p->mass_fraction_h) * 0.25;
    g.Rgas = p->spec_gas_const;
    g.dt_max = p->dt_max;
    g.g = p->gamma;
    g.gm1 = p->gamma - 1.0;
    g.g2 = p->gamma / (p->gamma - 1.0);
    g.gf = (float)p->gamma;
    g.gm1f = (float)p->gamma - 1.0f;
    g.g2f = (float)p->gamma / ((float)p->gamma - 1.0f);
    g.igm1f = 1.0f / g.gm1f;
    g.igm1 = 1.0 / g.gm1;
    g.idx = 1.0 / g.dx;

    c->stream = (hipStream_t)stream;  // NULL: the null (default) stream
    const size_t bytes = state_bytes(c);
    if (hipMalloc(&c->U, bytes) != hipSuccess || hipMalloc(&c->Ustar, bytes) != hipSuccess ||
        hipMalloc(&c->st, sizeof(tg::StepState)) != hipSuccess) {
        torchgpu_destroy(c);
        return TORCHGPU_ERR_HIP;
    }
    if (p->precision == TORCHGPU_MIXED &&
        (hipMalloc(&c->qf[0], bytes / 2) != hipSuccess || hipMalloc(&c->qf[1], bytes / 2) != hipSuccess)) {
        torchgpu_destroy(c);
        return TORCHGPU_ERR_HIP;
    }
    if (reset_clock(c) != TORCHGPU_OK) {
        torchgpu_destroy(c);
        return TORCHGPU_ERR_HIP;
    }
    *out = c;
    return TORCHGPU_OK;
}

void torchgpu_destroy(torchgpu_ctx* c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    if (c->U) (void)hipFree(c->U);
    if (c->Ustar) (void)hipFree(c->Ustar);
    if (c->st) (void)hipFree(c->st);
    if (c->qprint) (void)hipFree(c->qprint);
    if (c->pstate) (void)hipFree(c->pstate);
    if (c->qf[0]) (void)hipFree(c->qf[0]);
    if (c->qf[1]) (void)hipFree(c->qf[1]);
    if (c->heat) (void)hipFree(c->heat);
    delete c;
}

int torchgpu_last_error(const torchgpu_ctx* c, char* buf, size_t len) {
    if (!c || !buf || len == 0) return TORCHGPU_ERR_INVALID_ARG;
    std::snprintf(buf, len, "%s", c->err.c_str());
    return TORCHGPU_OK;
}

int torchgpu_state_ptr(torchgpu_ctx* c, double** U) {
    if (!c || !U) return TORCHGPU_ERR_INVALID_ARG;
    c->qvalid = false;  // the caller may write U through it
    *U = c->U;
    return TORCHGPU_OK;
}

int torchgpu_set_primitive(torchgpu_ctx* c, const double* Q, int on_device) {
    if (!c || !Q) return TORCHGPU_ERR_INVALID_ARG;
    HIP_TRY(c, hipMemcpyAsync(c->Ustar, Q, state_bytes(c), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, tg::launch_u_from_q(c->g, c->Ustar, c->U, c->stream));
    c->qvalid = c->qstar_valid = false;
    // a new set-up: GridCell::H starts at zero (GridCell.cpp:30-39)
    if (c->heat) HIP_TRY(c, tg::launch_zero(c->heat, 3 * c->g.ncell, c->stream));
    return seed_scale(c);
}

int torchgpu_set_conserved(torchgpu_ctx* c, const double* U, int on_device) {
    if (!c || !U) return TORCHGPU_ERR_INVALID_ARG;
    HIP_TRY(c, hipMemcpyAsync(c->U, U, state_bytes(c), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                              c->stream));
    c->qvalid = false;
    return seed_scale(c);
}

int torchgpu_get_conserved(torchgpu_ctx* c, double* U, int on_device) {
    if (!c || !U) return TORCHGPU_ERR_INVALID_ARG;
    HIP_TRY(c, hipMemcpyAsync(U, c->U, state_bytes(c), on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return TORCHGPU_OK;
}

int torchgpu_get_primitive(torchgpu_ctx* c, double* Q, int fix, int on_device) {
    if (!c || !Q) return TORCHGPU_ERR_INVALID_ARG;
    HIP_TRY(c, tg::launch_q_from_u(c->g, c->U, c->Ustar, fix, c->stream));
    c->qstar_valid = false;  // U* holds Q now
    HIP_TRY(c, hipMemcpyAsync(Q, c->Ustar, state_bytes(c), on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return TORCHGPU_OK;
}

int torchgpu_hydro_timestep(torchgpu_ctx* c, double dt_max, double* dt) {
    if (!c || !dt) return TORCHGPU_ERR_INVALID_ARG;
    // a private StepState slot: reuse the context's maxinv after saving it
    tg::StepState saved;
    HIP_TRY(c, hipMemcpyAsync(&saved, c->st, sizeof saved, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    tg::StepState tmp = saved;
    tmp.maxinv = 0ull;
    HIP_TRY(c, hipMemcpyAsync(c->st, &tmp, sizeof tmp, hipMemcpyHostToDevice, c->stream));
    tg::HydroGeom g = c->g;
    g.dt_max = dt_max;
    HIP_TRY(c, tg::launch_cfl(g, c->U, c->st, c->stream));
    HIP_TRY(c, hipMemcpyAsync(&tmp, c->st, sizeof tmp, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->st, &saved, sizeof saved, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // Hydro.cpp:78-88: tmin = min(dt_max, min_cells 0.5/inv_t) = min(dt_max, 0.5/max inv_t)
    double maxinv;
    std::memcpy(&maxinv, &tmp.maxinv, sizeof maxinv);
    double tmin = dt_max;
    if (maxinv > 0.0) { double t = 0.5 / maxinv; tmin = (t < tmin) ? t : tmin; }
    *dt = (dt_max < tmin) ? dt_max : tmin;
    return TORCHGPU_OK;
}

static int hydro_step_device(torchgpu_ctx* c, bool last = true) {
    // Torch::hydroStep: predictor U* = fix(U + dt/2 L(Q(U))), corrector
    // U = fix(U + dt L(Q(U*))); U is updated in place by the corrector.
    int rc = run_stage(c, 1, -1.0, last);
    if (rc) return rc;
    return run_stage(c, 2, -1.0);
}

int torchgpu_hydro_step(torchgpu_ctx* c, double dt) {
    if (!c) return TORCHGPU_ERR_INVALID_ARG;
    c->last_mode = 1;
    // an explicit dt: step_begin with use_cfl = 0 still honours tmax halting
    // bookkeeping but takes dt as given
    HIP_TRY(c, tg::launch_step_begin(c->st, 0, dt, c->stream));
    int rc = hydro_step_device(c);
    if (rc) return rc;
    HIP_TRY(c, tg::launch_step_end(c->st, c->stream));
    return TORCHGPU_OK;
}

int torchgpu_set_dt(torchgpu_ctx* c, double dt) {
    if (!c) return TORCHGPU_ERR_INVALID_ARG;
    HIP_TRY(c, tg::launch_step_begin(c->st, 0, dt, c->stream));
    return TORCHGPU_OK;
}

int torchgpu_hydro_stage(torchgpu_ctx* c, int stage) {
    if (!c || (stage != 1 && stage != 2)) return TORCHGPU_ERR_INVALID_ARG;
    return run_stage(c, stage, -1.0);
}

int torchgpu_hydro_stage_dt(torchgpu_ctx* c, int stage, double dtf) {
    if (!c || stage < 0 || stage > 2 || !(dtf >= 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    // stage 0: the single-stage sub-step (Torch::subStep with the hydro
    // component), written out of place, then the two state buffers swap roles
    return run_stage(c, stage, dtf);
}

int torchgpu_run(torchgpu_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return TORCHGPU_ERR_INVALID_ARG;
    c->last_mode = 1;
    for (int s = 0; s < nsteps; ++s) {
        HIP_TRY(c, tg::launch_step_begin(c->st, 1, 0.0, c->stream));
        int rc = hydro_step_device(c, s == nsteps - 1);
        if (rc) return rc;
        HIP_TRY(c, tg::launch_step_end(c->st, c->stream));
    }
    return TORCHGPU_OK;
}

int torchgpu_reset_clock(torchgpu_ctx* c) {
    if (!c) return TORCHGPU_ERR_INVALID_ARG;
    int rc = reset_clock(c);
    if (rc) return rc;
    return seed_scale(c);
}

int torchgpu_set_time(torchgpu_ctx* c, double t) {
    if (!c || !(t >= 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    HIP_TRY(c, tg::launch_set_time(c->st, t, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return TORCHGPU_OK;
}

int torchgpu_pause_at_checkpoints(torchgpu_ctx* c, int on) {
    if (!c) return TORCHGPU_ERR_INVALID_ARG;
    tg::StepState st;
    HIP_TRY(c, hipMemcpyAsync(&st, c->st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    st.pause_ckpt = on ? 1 : 0;
    HIP_TRY(c, hipMemcpyAsync(c->st, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->pause_ckpt = on != 0;
    return TORCHGPU_OK;
}

int torchgpu_checkpoint(torchgpu_ctx* c, int* count, int* due) {
    if (!c || !count || !due) return TORCHGPU_ERR_INVALID_ARG;
    tg::StepState st;
    HIP_TRY(c, hipMemcpyAsync(&st, c->st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *count = st.ckpt_count;
    *due = st.paused;
    if (st.paused) {
        HIP_TRY(c, tg::launch_resume(c->st, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return TORCHGPU_OK;
}

int torchgpu_get_print_state(torchgpu_ctx* c, double* Q, int on_device) {
    if (!c || !Q) return TORCHGPU_ERR_INVALID_ARG;
    int rc = print_state(c);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(Q, c->qprint, state_bytes(c), on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return TORCHGPU_OK;
}

int torchgpu_heating_track(torchgpu_ctx* c, int on) {
    if (!c) return TORCHGPU_ERR_INVALID_ARG;
    if (!on) {
        if (c->heat) {
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            (void)hipFree(c->heat);
            c->heat = nullptr;
        }
        return TORCHGPU_OK;
    }
    if (!c->heat) {
        HIP_TRY(c, hipMalloc(&c->heat, 3 * sizeof(double) * (size_t)c->g.ncell));
        HIP_TRY(c, tg::launch_zero(c->heat, 3 * c->g.ncell, c->stream));
    }
    return TORCHGPU_OK;
}

int torchgpu_synchronize(torchgpu_ctx* c) {
    if (!c) return TORCHGPU_ERR_INVALID_ARG;
    return check_device_errors(c);
}

int torchgpu_get_clock(torchgpu_ctx* c, double* time, double* last_dt, int* steps) {
    if (!c) return TORCHGPU_ERR_INVALID_ARG;
    tg::StepState st;
    HIP_TRY(c, hipMemcpyAsync(&st, c->st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (time) *time = st.time;
    if (last_dt) *last_dt = st.dt;
    if (steps) *steps = st.step;
    return TORCHGPU_OK;
}

int torchgpu_is_quitting(torchgpu_ctx* c, int* quitting) {
    if (!c || !quitting) return TORCHGPU_ERR_INVALID_ARG;
    tg::StepState st;
    HIP_TRY(c, hipMemcpyAsync(&st, c->st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *quitting = st.quitting;
    return TORCHGPU_OK;
}

}  // extern "C"
