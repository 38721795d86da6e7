// torchgpu_internal.hpp -- device-side parameter blocks, the per-run step
// state and the launcher entry points shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace tg {

// Grid + hydro parameters, passed by value to every hydro kernel.
struct HydroGeom {
    int nx, ny, nz;
    int ntx, nty, ntz;      // 8^3 tiles per dimension
    int ntiles;
    int bc_l[3], bc_r[3];
    int limiter;
    int spatial_order;
    int addr64;             // mixed kernel: 64-bit addressing even where 32-bit offsets fit
    long long ncell;        // nx*ny*nz, SoA stride
    double dx;
    double kflux;           // (area / vol) of a Cartesian face, Hydro.cpp:129-130
    double dfloor, pfloor, tfloor;
    double X;               // mass fraction of hydrogen
    double mu_c;            // (1.0 - X) * 0.25
    double Rgas;            // specific gas constant (code units)
    double dt_max;
    double g, gm1, g2;      // gamma, gamma - 1.0, gamma / (gamma - 1.0)
    float gf, gm1f, g2f;    // the same rounded to float (mixed path)
    float igm1f;            // 1 / gm1f (mixed path)
    double igm1, idx;       // 1 / gm1, 1 / dx (mixed path)
};

// A star at a cell centre of a 3-D grid (cell units: cells at index + 0.5,
// the star at its cell + 0.5), with its wind radius and photon rate, for the
// star-on thermo operator (Star::initialise without face snapping,
// Star.cpp:17-21; Fluid::initialiseGrid, Fluid.cpp:27-57).
struct StarGeom {
    int on;
    int s[3];               // the star's cell
    int n[3];               // nx, ny, nz
    int wind_r;             // Star.wind_radius_in_cells
    double dx;              // cell width (code units)
    double photon_rate;     // code units
    double pi;              // Constants::pi
    double hydrogen_mass;   // code units
};

// One Chebyshev shell of cells around the star as up to six faces clipped to
// the grid: face q fixes coordinate `axis[q]` at `fixed[q]` and spans the two
// other axes (lower index first) from lo0/lo1, w0 cells along the first.
struct ShellFaces {
    int nf;
    int axis[6], fixed[6], lo0[6], lo1[6], w0[6];
    long long start[7];     // prefix sums of the face sizes
};

// Device-resident state of a time-stepping run (one per context).
struct StepState {
    double dt;              // dt of the step in flight
    double time;            // simulation time (code units)
    double dt_max;
    double tmax;
    unsigned long long maxinv;  // bit pattern of max_cells inv_t (>= 0)
    int first;              // next step is the reference's first step
    int ckpt_count;
    int ncheckpoints;
    int halted;             // time >= tmax, or a step raised an error: further steps are no-ops
    unsigned int err;       // error bits (ERR_*) raised during step err_step
    int err_step1;          // 1 + step index of the first error (0: none)
    int step;               // steps completed
    int debug;              // Integration.debug: quit when the hydro dt is too small (Torch.cpp:321-332)
    // mixed path: float bit patterns of the largest density and pressure
    // the state scale is taken from, and of those the running predictor
    // records for the next one (hydro_mixed.hip)
    unsigned int qmax[2];
    unsigned int qmax_next[2];
    // Torch::run prints a snapshot when a new checkpoint is due
    // (Torch.cpp:268-280): with pause_ckpt set, the run stops before such a
    // step (paused, halted) until the host has printed and resumes it
    int pause_ckpt;
    int paused;
    // Torch::m_isQuitting: set by the step whose hydro time step had
    // 100 dt_hydro / tmax <= 1e-6 with debug on; the steps after it halt
    int quitting;
    // the step in flight is the last before a print can happen (it reaches
    // a new checkpoint or tmax, or quits): the mixed predictor then stores
    // U*, whose QfromU Torch::run prints after a hydro step
    int keep_ustar;
};

// Negative dt arguments of the stage and cooling kernels take the dt of the
// step in flight from the device (StepState::dt), so that a time loop runs
// without host round trips: DT_STEP_RULE = dt/2 for a predictor, dt for a
// corrector (Torch::hydroStep); DT_HALF / DT_FULL = 0.5 dt / dt (the
// sub-steps of Torch::fullStep).  The kernels given a code skip their work
// once the run has halted.
constexpr double DT_STEP_RULE = -1.0, DT_HALF = -2.0, DT_FULL = -3.0;

// Error bits in the order the reference would meet them inside one step;
// the lowest set bit of the first failing step is the error reported.
enum : unsigned int {
    ERR_CHECK_VALUES = 1u,    // Torch::checkValues("hydro before") throw (Torch.cpp:420-443)
    ERR_RIEMANN_NAN_1 = 2u,   // RiemannSolver::solve: HII flux is NaN, predictor (Riemann.cpp:95-103)
    ERR_FIX_SOLUTION_1 = 4u,  // Fluid::fixSolution throw, predictor (Fluid.cpp:91-92,128-134)
    ERR_RIEMANN_NAN_2 = 8u,   // ... corrector
    ERR_FIX_SOLUTION_2 = 16u, // ... corrector
    ERR_THERMO_NSTEPS = 32u,  // subcycle count beyond the int range of Thermodynamics.cpp:307
    ERR_THERMO_FIX = 64u,     // fixSolution throw after a thermo or radiation sub-step
    ERR_RAD_CHECK = 128u,     // checkValues("Radiationbefore") throw (Torch.cpp:353,420-443)
    ERR_RAD_NAN = 256u,       // "heating rate in radiation module is NaN" (Radiation.cpp:284-295)
    ERR_RAD_ITER = 512u,      // "implicit method not converging" (Radiation.cpp:436-455)
    ERR_RAD_DT = 1024u,       // "Radiation::calculateTimeStep(): invalid timestep" (Radiation.cpp:346-347)
};

// Device view of a grid context (api.hip), for operators implemented in
// other translation units.
struct CtxView {
    double* U;
    double* Ustar;
    StepState* st;
    hipStream_t stream;
    const HydroGeom* g;
    // GridCell::H kept across steps while heating output is tracked
    // (torchgpu_heating_track), else null: [3][n] TOT, RHII, EUVH
    double* heat;
};

// The three heating terms a grid keeps between Thermodynamics::
// fillHeatingArrays calls (HEAT_TOT set by the thermo integrate, HEAT_RHII
// and HEAT_EUVH by the radiation pre-step); the other eight are recomputed
// from Q by each fill.
enum { HEAT_TOT = 0, HEAT_RHII = 1, HEAT_EUVH = 2 };

}  // namespace tg
typedef struct torchgpu_ctx torchgpu_ctx;
typedef struct torchgpu_radiation torchgpu_radiation;
namespace tg {
CtxView ctx_view(torchgpu_ctx* c);

// Launchers (hydro.hip).  All enqueue on `stream` and return hipError_t.
// cerr (may be null): a cooling operator's error word, whose subcycle-count
// overflow is moved into st (and the word cleared).
hipError_t launch_fix_solution(const HydroGeom& g, double* U, const double* dpre, double dtf, StepState* st,
                               hipStream_t stream, unsigned int* cerr = nullptr);
// FP64 stage kernel (hydro.hip).  dtf < 0: the stage takes its dt from st
// (dt/2 predictor, dt corrector).
hipError_t launch_hydro_stage(const HydroGeom& g, int solver, int stage, const double* Usrc, const double* Ubase,
                              double* Udst, StepState* st, hipStream_t stream, double dtf = -1.0);
// Mixed-precision stage kernel (hydro_mixed.hip): FP32 reconstruction and
// Riemann flux, FP64 accumulation, pressure recovery and CFL.  It reads the
// float primitives qin of its input state and writes those of its output
// (qout, with fixPrimitives and the "hydro before" check if qout_fix) for the
// next stage; launch_q_float produces them for any other state.
hipError_t launch_mixed_scale_seed(const HydroGeom& g, const double* U, StepState* st, hipStream_t stream);
hipError_t launch_q_float(const HydroGeom& g, const double* U, float* q, int fix, hipStream_t stream);
// launch_fix_solution fused with launch_q_float(U, qout, fix 1) on the result.
hipError_t launch_fix_solution_qf(const HydroGeom& g, double* U, const double* rate, double dtf, StepState* st,
                                  hipStream_t stream, unsigned int* cerr, float* qout);
hipError_t launch_hydro_stage_mixed(const HydroGeom& g, int solver, int stage, const float* qin, const double* Ubase,
                                    double* Udst, float* qout, int qout_fix, StepState* st, hipStream_t stream,
                                    double dtf = -1.0, int ustar = 1);
// The grid state was changed by another operator (the float primitives of
// the mixed path are recomputed before the next stage).
void ctx_state_changed(torchgpu_ctx* c);
// Mixed precision: the float Q buffer that must hold Q(U) for the next stage
// (NULL for FP64 contexts), and the mark that it now does.
float* ctx_qf_for_state(torchgpu_ctx* c);
void ctx_qf_written(torchgpu_ctx* c);
// One hydro stage on the context's buffers (api.hip): mode 1 predictor,
// 2 corrector, 0 single-stage sub-step; dtf as above.
int ctx_run_stage(torchgpu_ctx* c, int mode, double dtf);
// The context's Strang alternation counter (Torch::stepCounter).
int& ctx_step_counter(torchgpu_ctx* c);
// The device step count at which ctx_step_counter was last brought up to
// date (the device loop launches steps that may not run: a checkpoint pause
// or tmax halts them, and only the steps that ran turn the counter).
int& ctx_counter_step(torchgpu_ctx* c);
// Bookkeeping of the state Torch::run prints (see torchgpu_get_print_state):
// the mode of the last run, and (when pausing at checkpoints) the buffer
// that keeps the input U of an operator-split step's last sub-step (mode
// 2), or the printed Q itself (mode 3: runs with radiation, whose sub-step
// leaves its HII fraction in Fluid::Q).
void ctx_set_print_mode(torchgpu_ctx* c, int mode);
double* ctx_print_buffer(torchgpu_ctx* c);
hipError_t launch_cfl(const HydroGeom& g, const double* U, StepState* st, hipStream_t stream);
// p[0 .. n) = 0 with a kernel (no memset node if a caller captures a graph).
hipError_t launch_zero(double* p, long long n, hipStream_t stream);
// Fluid::Q as Torch::run prints it (see torchgpu_get_print_state), in a
// device buffer of the context; null on error (message in the context).
const double* ctx_print_state_dev(torchgpu_ctx* c);
// thermo_dt_bits: the bit pattern of the cooling operator's minimum
// |E/RATE| (a Torch::fullStep with thermodynamics), or null; rad_dt_bits:
// that of Radiation::calculateTimeStep's minimum, or null.
hipError_t launch_step_begin(StepState* st, int use_cfl, double dt_override, hipStream_t stream,
                             const unsigned long long* thermo_dt_bits = nullptr,
                             const unsigned long long* rad_dt_bits = nullptr);
hipError_t launch_step_end(StepState* st, hipStream_t stream);
// Clock control for Torch::run's checkpoint printing: set the time (and the
// checkpoint count Checkpointer::update gives for it), resume a paused run.
hipError_t launch_set_time(StepState* st, double t, hipStream_t stream);
// The column-density ray trace (Thermodynamics::rayTrace): one launch per
// shell, den/col/dcol [n] device arrays (den floored at den_floor, as
// fixPrimitives does; -HUGE_VAL: as given); the star's geometry per cell.
hipError_t launch_thermo_ray_trace(const StarGeom& s, const double* den, double den_floor, double* col, double* dcol,
                                   double* xface, hipStream_t stream);
// doubles of the x-face buffers of launch_thermo_ray_trace
long long xface_doubles(const StarGeom& s);
// The cells of Chebyshev shell k around the star, and the largest k.
void shell_faces(const StarGeom& s, int k, ShellFaces& f);
int shell_max(const StarGeom& s);
hipError_t launch_star_geometry(const StarGeom& s, long long* nb, double* w, double* ds, double* shell_vol,
                                hipStream_t stream);
// Pass-through of a halted single-stage sub-step (Usrc -> Udst, and the
// float primitives qin -> qout when qin is given); a no-op while running.
hipError_t launch_halted_copy(const HydroGeom& g, const StepState* st, const double* Usrc, double* Udst,
                              const float* qin, float* qout, hipStream_t stream);
hipError_t launch_copy_if_running(const StepState* st, const double* src, double* dst, long long n,
                                  hipStream_t stream);
hipError_t launch_resume(StepState* st, hipStream_t stream);
hipError_t launch_q_from_u(const HydroGeom& g, const double* U, double* Q, int fix, hipStream_t stream);
// Q = fixPrimitives(QfromU(U)) unless the run has halted.
hipError_t launch_q_from_u_if_running(const HydroGeom& g, const StepState* st, const double* U, double* Q,
                                      hipStream_t stream);
hipError_t launch_u_from_q(const HydroGeom& g, double* Q, double* U, hipStream_t stream);
// The radiation operator's pieces of Torch::fullStep for the device loop
// (radiation.hip).  dt of rad_substep_dev: explicit, or a DT_* code.
hipStream_t rad_stream(const torchgpu_radiation* r);
const unsigned long long* rad_dt_bits(const torchgpu_radiation* r);
int rad_pre_step_dev(torchgpu_radiation* r, int gated);
int rad_timestep_dev(torchgpu_radiation* r, double dt_max, int gated);
int rad_once_commit(torchgpu_radiation* r, int gated);
int rad_substep_dev(torchgpu_radiation* r, double dt, int has_heat);
int rad_print_hii(torchgpu_radiation* r, double* P);

}  // namespace tg
