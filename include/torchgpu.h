/*
 * torchgpu.h -- C ABI of the MI355X (gfx950) hydro + thermodynamics step of
 * TORCH (hgsteggles/TORCH).
 *
 * The library replaces, for one whole grid on one GPU, the reference's
 * operator-split step between the driver and the grid state:
 *
 *   reference interface                                   entry point here
 *   ---------------------------------------------------   ------------------------------
 *   Torch::initialise tail: Fluid::fixPrimitives +        torchgpu_set_primitive
 *     Fluid::globalUfromQ   (src/Torch/Torch.cpp:117-119)
 *   Fluid::globalQfromU / fixPrimitives                   torchgpu_get_primitive
 *     (src/Fluid/Fluid.cpp:138-150,162-165)
 *   GridCell::U state read/write (src/Fluid/GridCell.hpp:43-105) torchgpu_set_conserved /
 *                                                              torchgpu_get_conserved
 *   Hydrodynamics::calculateTimeStep                      torchgpu_hydro_timestep
 *     (src/Integrators/Hydro.cpp:77-89)
 *   Torch::hydroStep(dt, true): Hydrodynamics::integrate  torchgpu_hydro_step
 *     + updateSourceTerms + Fluid::advSolution/fixSolution,
 *     twice (src/Torch/Torch.cpp:366-387)
 *   Torch::run / fullStep / calculateTimeStep for a       torchgpu_run
 *     hydro-only run (src/Torch/Torch.cpp:240-302,304-339,389-404)
 *   Thermodynamics::preTimeStepCalculations + integrate   torchgpu_thermo_cells
 *     + updateSourceTerms on independent cells
 *     (src/Integrators/Thermodynamics.cpp:218-337,478-485)
 *   Thermodynamics::calculateTimeStep / the thermo part   torchgpu_thermo_grid_timestep /
 *     of Torch::subStep on the grid state                   torchgpu_thermo_grid_substep
 *     (Thermodynamics.cpp:463-485, Torch.cpp:352-364)
 *   Torch::run / fullStep / calculateTimeStep with        torchgpu_run_split,
 *     hydrodynamics and thermodynamics and/or radiation     torchgpu_run_full
 *     (Torch.cpp:240-418)
 *   Torch::run's Checkpointer test and the state it       torchgpu_pause_at_checkpoints /
 *     prints (Torch.cpp:268-277, Checkpointer.cpp:14-30)     torchgpu_checkpoint /
 *                                                            torchgpu_get_print_state
 *   Torch::initialise start time from initial conditions  torchgpu_set_time
 *     (Torch.cpp:100-103)
 *   DataPrinter::print2D, DataReader (src/IO/)            torchgpu_snapshot_*
 *   Radiation::initField / preTimeStepCalculations /      torchgpu_radiation_*
 *     calculateTimeStep / integrate / updateSourceTerms
 *     (src/Integrators/Radiation.cpp:78-680)
 *   Star-on thermo preTimeStep: Thermodynamics::rayTrace  torchgpu_cooling_set_star,
 *     (Thermodynamics.cpp:218-262,339-417), star geometry   torchgpu_star_geometry,
 *     (Grid.cpp:922-951, Radiation.cpp:78-138)              torchgpu_thermo_grid_inputs
 *   MLP surrogate of the subcycled update (no reference   torchgpu_mlp_*,
 *     counterpart; FP64 integrate remains the fallback)     torchgpu_thermo_cells_surrogate
 *
 * Conventions
 *   - State arrays are structure-of-arrays float64 [7][nz][ny][nx], x fastest,
 *     variables in the reference's UID order DEN, PRE, HII, ADV, VEL0..2
 *     (src/Torch/Common.hpp:23-25); U holds conserved and Q primitive
 *     variables.  All values are in code units (Converter, src/Torch/Converter.cpp).
 *   - Pointers flagged `on_device` are device (HBM) addresses; otherwise host.
 *   - Every call returns TORCHGPU_OK or an error code; torchgpu_last_error()
 *     gives the message (the reference's exception text where one exists).
 *   - Work is enqueued on the context's stream.  Calls returning host values
 *     synchronize it.  A context is not thread-safe.
 */
#ifndef TORCHGPU_H_
#define TORCHGPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TORCHGPU_OK = 0,
    TORCHGPU_ERR_INVALID_ARG = 1,
    TORCHGPU_ERR_HIP = 2,
    TORCHGPU_ERR_RIEMANN_NAN = 3,    /* "RiemannSolver::solve: HII flux is NaN" (Riemann.cpp:95-103,129-139) */
    TORCHGPU_ERR_FIX_SOLUTION = 4,   /* Fluid::fixSolution throw (Fluid.cpp:91-92,128-134) */
    TORCHGPU_ERR_CHECK_VALUES = 5,   /* Torch::checkValues throw (Torch.cpp:420-443) */
    TORCHGPU_ERR_INVALID_ORDER = 6,  /* Hydrodynamics::calcFluxes invalid order (Hydro.cpp:110-112) */
    TORCHGPU_ERR_UNKNOWN_SOLVER = 7, /* RiemannSolverFactory::create unknown type (Riemann.cpp:256) */
    TORCHGPU_ERR_UNKNOWN_LIMITER = 8,/* SlopeLimiterFactory::create unknown type (SlopeLimiter.cpp:73-74) */
    TORCHGPU_ERR_NSTEPS_RANGE = 9,   /* subcycle count beyond int range (Thermodynamics.cpp:307) */
    TORCHGPU_ERR_NOT_TRAINED = 10,   /* surrogate used before weights were set */
    TORCHGPU_ERR_RADIATION = 11      /* Radiation throws: NaN heating rate, invalid timestep, implicit
                                        HII iteration not converging, "Radiationbefore produced an error."
                                        (Radiation.cpp:284-295,346-347,436-455; Torch.cpp:420-443) */
};

/* Boundary conditions, Condition enum order of src/Torch/Common.hpp:27. */
enum { TORCHGPU_BC_FREE = 0, TORCHGPU_BC_REFLECTING = 1, TORCHGPU_BC_OUTFLOW = 2,
       TORCHGPU_BC_INFLOW = 3, TORCHGPU_BC_PERIODIC = 4 };
/* Riemann solvers, RiemannSolverFactory::create strings "HLL", "HLLC",
 * "RotatedHLLC" ("default"), src/Integrators/Riemann.cpp:246-257. */
enum { TORCHGPU_SOLVER_HLL = 0, TORCHGPU_SOLVER_HLLC = 1, TORCHGPU_SOLVER_ROTATED_HLLC = 2 };
/* Slope limiters, SlopeLimiterFactory::create strings
 * "monotonised_central", "superbee", "minmod", "maxmod", "leer", "ospre",
 * "albada" ("default"), src/Integrators/SlopeLimiter.cpp:56-75; the limiters
 * themselves are SlopeLimiter.cpp:7-54. */
enum { TORCHGPU_LIM_MC = 0, TORCHGPU_LIM_SUPERBEE = 1, TORCHGPU_LIM_MINMOD = 2, TORCHGPU_LIM_MAXMOD = 3,
       TORCHGPU_LIM_LEER = 4, TORCHGPU_LIM_OSPRE = 5, TORCHGPU_LIM_ALBADA = 6 };
/* Precision of the hydro flux path. */
enum { TORCHGPU_FP64 = 0, TORCHGPU_MIXED = 1 };

/* Parameters in code units, as TorchParameters holds them after
 * TorchParameters::initialise (src/Torch/Parameters.cpp:9-27).  Cartesian
 * 3-D geometry; cell width dx = side_length / nx for every dimension
 * (src/Fluid/Grid.cpp:471-472). */
typedef struct torchgpu_params {
    int nx, ny, nz;
    int spatial_order;        /* Integration.spatial_order: 0 or 1 */
    int bc_left[3], bc_right[3];
    int riemann_solver;       /* TORCHGPU_SOLVER_* */
    int slope_limiter;        /* TORCHGPU_LIM_* */
    int precision;            /* TORCHGPU_FP64 or TORCHGPU_MIXED */
    int ncheckpoints;         /* Integration.ncheckpoints */
    int debug;                /* Integration.debug (Parameters.hpp:50, default true): the device loops
                                 end after a step whose hydro dt has 100 dt_hydro / tmax <= 1e-6
                                 (Torch::calculateTimeStep, Torch.cpp:321-332) */
    double side_length;       /* Grid.side_length */
    double gamma;             /* Hydrodynamics.gamma */
    double dfloor;            /* density floor */
    double pfloor;            /* 0.1 * R * dfloor * tfloor (Parameters.cpp:18) */
    double tfloor;            /* Hydrodynamics.temperature_floor */
    double mass_fraction_h;   /* Radiation.mass_fraction_hydrogen */
    double spec_gas_const;    /* Constants::specificGasConstant */
    double tmax;              /* simulation time */
    double dt_max;            /* tmax / 10 (Parameters.cpp:11) */
} torchgpu_params;

typedef struct torchgpu_ctx torchgpu_ctx;

/* Context: owns the device state (U, the predictor buffer, step state) on
 * the current HIP device.  `stream` is the hipStream_t all of its work is
 * enqueued on (NULL: the null stream). */
int torchgpu_create(const torchgpu_params* params, void* stream, torchgpu_ctx** out);
void torchgpu_destroy(torchgpu_ctx* ctx);
int torchgpu_last_error(const torchgpu_ctx* ctx, char* buf, size_t len);
/* Device address of the conserved state U ([7][nz][ny][nx] float64).
 * Mixed precision: the float state is scaled by powers of two taken from
 * the largest density and pressure, which set_primitive, set_conserved and
 * reset_clock measure and every predictor launch records for the next one;
 * after writing U through this pointer by more than ~1e14x, call
 * torchgpu_reset_clock (or set_conserved) to re-measure. */
int torchgpu_state_ptr(torchgpu_ctx* ctx, double** U);

/* Torch::initialise tail (Torch.cpp:117-119): Q is fixed by fixPrimitives
 * (with U still zero) and U = UfromQ(Q).  Q is not modified on the host. */
int torchgpu_set_primitive(torchgpu_ctx* ctx, const double* Q, int on_device);
int torchgpu_set_conserved(torchgpu_ctx* ctx, const double* U, int on_device);
int torchgpu_get_conserved(torchgpu_ctx* ctx, double* U, int on_device);
/* Q = QfromU(U), then fixPrimitives if `fix` (Fluid.cpp:138-165). */
int torchgpu_get_primitive(torchgpu_ctx* ctx, double* Q, int fix, int on_device);

/* Hydrodynamics::calculateTimeStep(dt_max) on fixPrimitives(QfromU(U)). */
int torchgpu_hydro_timestep(torchgpu_ctx* ctx, double dt_max, double* dt);
/* Torch::hydroStep(dt, true).  Asynchronous; errors surface at the next
 * synchronizing call. */
int torchgpu_hydro_step(torchgpu_ctx* ctx, double dt);
/* One stage of Torch::hydroStep: Hydrodynamics::integrate + updateSourceTerms
 * on Q = QfromU(U) (fixPrimitives'd when stage == 1), then
 * Fluid::advSolution(h * dt) + fixSolution applied to U, written to the
 * predictor buffer (stage 1, h = 1/2) or in place to U (stage 2, h = 1;
 * Q from the predictor buffer, CFL reduction of the next step fused).
 * Exposed for per-kernel timing and for drivers that sequence the stages
 * themselves (Torch.cpp:374-386).  dt is the last dt set by
 * torchgpu_hydro_step / torchgpu_run / torchgpu_set_dt. */
int torchgpu_hydro_stage(torchgpu_ctx* ctx, int stage);
/* Sets the dt used by torchgpu_hydro_stage. */
int torchgpu_set_dt(torchgpu_ctx* ctx, double dt);
/* As torchgpu_hydro_stage with the scaled dt given (dtf = dt/2.0 for the
 * predictor, dt for the corrector); stage 0 is the single-stage sub-step of
 * an operator-split step (Torch::subStep with Hydrodynamics,
 * Torch.cpp:352-364): U <- fixSolution(U + dtf * L(fixPrimitives(QfromU(U)))).
 * The state pointer of torchgpu_state_ptr may change after stage 0. */
int torchgpu_hydro_stage_dt(torchgpu_ctx* ctx, int stage, double dtf);
/* Up to `nsteps` iterations of the Torch::run loop for a hydro-only run,
 * with dt chosen on the device (first step dt_max * 1e-20, then the CFL
 * min-reduction fused into the corrector stage, clipped to the next
 * checkpoint).  Asynchronous. */
int torchgpu_run(torchgpu_ctx* ctx, int nsteps);
/* Restart the Torch::run bookkeeping (time 0, first-step rule, checkpoints);
 * mixed precision: also re-measures the state scale.  Synchronises. */
int torchgpu_reset_clock(torchgpu_ctx* ctx);
/* Waits for the stream and reports the first device-side error. */
int torchgpu_synchronize(torchgpu_ctx* ctx);
/* Simulation time, dt of the last step and number of completed steps. */
int torchgpu_get_clock(torchgpu_ctx* ctx, double* time, double* last_dt, int* steps);
/* Torch::isQuitting (src/Torch/Torch.hpp:47): set, with params.debug, by the
 * step whose hydro time step gave 100 dt_hydro / tmax <= 1e-6
 * ("Integration deltas are too small.", Torch.cpp:321-332); that step runs,
 * later steps of torchgpu_run / run_split / run_full do nothing (the
 * reference's Torch::run loop ends).  Cleared by torchgpu_reset_clock.
 * Synchronises. */
int torchgpu_is_quitting(torchgpu_ctx* ctx, int* quitting);
/* Torch::initialise's start time from an initial-conditions snapshot
 * (src/Torch/Torch.cpp:100-103, code units; t >= 0).  Synchronises. */
int torchgpu_set_time(torchgpu_ctx* ctx, double t);
/* With `on`, torchgpu_run, torchgpu_run_split and torchgpu_run_full stop at
 * the end of a step that reaches a checkpoint (the Checkpointer test of
 * Torch::run, src/Torch/Torch.cpp:268-277, src/IO/Checkpointer.cpp:14-30): later steps
 * of the same call do nothing until torchgpu_checkpoint resumes the run.
 * Also keeps the state Torch::run prints at that point.  Synchronises. */
int torchgpu_pause_at_checkpoints(torchgpu_ctx* ctx, int on);
/* The Checkpointer's count and whether the run is paused at a checkpoint
 * (`due`); a paused run is resumed.  Synchronises. */
int torchgpu_checkpoint(torchgpu_ctx* ctx, int* count, int* due);
/* Fluid::Q as Torch::run prints it (Torch.cpp:271-276, DataPrinter::print2D):
 * after a hydro-only step Q(U*) without fixPrimitives (Torch.cpp:366-387);
 * after a full step with thermodynamics fixPrimitives(Q(input of the last
 * substep)) (Torch.cpp:341-364), which needs pause_at_checkpoints on; before
 * any step fixPrimitives(Q(U)).  Q is [7][nz][ny][nx] code units, host
 * unless `on_device`.  Synchronises. */
int torchgpu_get_print_state(torchgpu_ctx* ctx, double* Q, int on_device);


/* ------------------------------------------------------------------------ */
/* Snapshots in the reference's data2D text format (host-only)                */
/* ------------------------------------------------------------------------ */
/* DataPrinter::print2D (src/IO/DataPrinter.cpp:139-174): time t and the
 * primitive state Q (host [7][nz][ny][nx], code units) written as text --
 * t (s), nx, ny, nz, then per cell (x fastest) the nd cell-centre
 * coordinates (cm), density, pressure, HII fraction and nd velocities (cgs),
 * tab separated, "%.10e" -- gzip-compressed when `gz` (as the reference's
 * .txt.gz).  nd = 1 if ny == 1, 2 if nz == 1, else 3; dx = side_length / nx
 * (code units); scales as torchgpu_units. */
int torchgpu_snapshot_write(const char* path, int gz, double t, const double* Q, int nx, int ny, int nz,
                            double side_length, double density_scale, double pressure_scale, double time_scale);
/* DataReader::readDataParameters (src/IO/DataReader.cpp:17-68): time (s),
 * cell counts, nd and side length (cm: dx * nx, dx from the first two
 * cells' coordinates) of a snapshot, plain or gzip text. */
int torchgpu_snapshot_header(const char* path, double* t, int* nx, int* ny, int* nz, int* nd, double* side_length);
/* DataReader::readGrid (src/IO/DataReader.cpp:70-125): density, pressure,
 * HII fraction and the nd velocities (cgs, as stored) of the first
 * nx*ny*nz cells listed, each placed by its coordinates / dx into Q (host
 * [7][nz][ny][nx]); cells outside the grid are skipped, other entries of Q
 * are left as they are. */
int torchgpu_snapshot_read(const char* path, double* Q, int nx, int ny, int nz);
/* DataReader::patchGrid (src/IO/DataReader.cpp:126-231), on host primitives
 * Q [7][nz][ny][nx] as they stand before Torch::toCodeUnits: each patch
 * cell lands in grid cell (int)(coordinate / grid_dx + offset); the covered
 * cells are zeroed, then receive the sum of their patch cells' density,
 * pressure, HII fraction and velocities divided by r = ratio^nd, ratio =
 * grid_dx / patch dx.  grid_dx is the grid's cell width as the reference's
 * Grid holds it (code units -- the reference divides the patch's cgs
 * coordinates by it).  The ratio and the covered cell counts must be
 * integers and nd must match. */
int torchgpu_snapshot_patch(const char* path, double* Q, int nx, int ny, int nz, double grid_dx, int ox, int oy,
                            int oz);

/* ------------------------------------------------------------------------ */
/* Thermodynamics (cooling) operator                                         */
/* ------------------------------------------------------------------------ */

/* Constants in code units and the ThermoParameters switches, as
 * Thermodynamics::initialise holds them (Thermodynamics.cpp:24-80) after
 * Constants::initialise (Constants.cpp:27-40). */
typedef struct torchgpu_thermo_params {
    double thermo_hii_switch, heating_amplification, mass_fraction_h, gamma, pfloor;
    double hydrogen_mass, spec_gas_const, dust_xsec, pi;
    double z0, T1, T2, T3, T4, imlc, nmlc, ciec_minT, ciec, cxhi_damp, n0, nmc;
    double fuvh_a, fuvh_b, fuvh_c, irh_a, irh_b, crh;
    double spl_min_slope, spl_max_slope;
    double spl_x[26], spl_y[26], spl_y2[26];
    int subcycling;
    int pad;
} torchgpu_thermo_params;

/* Converter::toCodeUnits / fromCodeUnits (Converter.cpp:27-63) for the
 * scales of Integration.{density,pressure,time}_scale. */
int torchgpu_units(double density_scale, double pressure_scale, double time_scale, double value,
                   double mass_index, double length_index, double time_index, int to_code, double* out);
/* Fills `t` from the Lua-level parameters (cgs floors, code scales). */
int torchgpu_thermo_params_init(torchgpu_thermo_params* t, double density_scale, double pressure_scale,
                                double time_scale, double thermo_hii_switch, double heating_amplification,
                                double mass_fraction_h, double gamma, double density_floor_cgs,
                                double temperature_floor, int subcycling);

typedef struct torchgpu_cooling torchgpu_cooling;

/* A cooling operator on the current device (spline tables, work counters and
 * the surrogate network live here). */
int torchgpu_cooling_create(const torchgpu_thermo_params* t, void* stream, torchgpu_cooling** out);
void torchgpu_cooling_destroy(torchgpu_cooling* c);
int torchgpu_cooling_last_error(const torchgpu_cooling* c, char* buf, size_t len);

/* Thermodynamics::preTimeStepCalculations + integrate on n independent
 * non-wind cells (Thermodynamics.cpp:218-337), FP64, subcycled.
 * in  (device, float64 [9][n]): den, pre, hii, adv, E (= U[PRE]), T_min,
 *      COL_DEN, F_FUV, dt -- code units.
 * out (device, float64 [5][n]): T[HEAT], T[RATE] after preTimeStep, T[RATE]
 *      after integrate, number of subcycle rate evaluations, subcycled
 *      pressure.
 * total_substeps (host, may be NULL) receives the sum of subcycle counts
 * (this synchronizes). */
int torchgpu_thermo_cells(torchgpu_cooling* c, int64_t n, const double* in, double* out, int64_t* total_substeps);
/* torchgpu_thermo_cells without allocating device memory.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * longest-first work order that torchgpu_thermo_cells keeps in buffers the
 * handle grows lives here in one fixed device array of the library (9 MiB per
 * GPU, held until the process ends), so n is limited to 1048576 cells (more:
 * TORCHGPU_ERR_INVALID_ARG, nothing launched).  Results equal
 * torchgpu_thermo_cells'.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams or threads) race.  The handle's device
 * must be current. */
int torchgpu_thermo_cells_noalloc(torchgpu_cooling* c, int64_t n, const double* in, double* out,
                                  int64_t* total_substeps);

/* Thermodynamics on the grid state of ctx (operator-split step).  Per-cell
 * inputs come from U (Q = fixPrimitives(QfromU(U)), E = U[PRE]); T_min,
 * col_den (T[COL_DEN]) and F_FUV are device arrays [n] or NULL (100 K, 0,
 * 0).  `excluded` is the flat index of the star's cell (CausalWind,
 * Fluid.cpp:41-54) or -1. */
/* Thermodynamics::calculateTimeStep (Thermodynamics.cpp:463-476) after
 * preTimeStepCalculations on the current state. */
int torchgpu_thermo_grid_timestep(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                                  const double* F_FUV, int64_t excluded, double dt_max, double* dt);
/* The thermo part of Torch::subStep (Torch.cpp:352-364): preTimeStep +
 * integrate(dt) (subcycled FP64, or the surrogate when `surrogate`) +
 * updateSourceTerms + Fluid::advSolution(dt) + fixSolution. */
int torchgpu_thermo_grid_substep(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                                 const double* F_FUV, int64_t excluded, double dt, int surrogate);
/* Up to `nsteps` iterations of the Torch::run loop for a run with
 * hydrodynamics and thermodynamics (Torch.cpp:240-302): each a
 * Torch::fullStep (Torch.cpp:389-418) with dt = min(dt_nextCheckpoint,
 * calculateTimeStep()) -- dt_max * 1e-20 first, then min(dt_max, hydro CFL,
 * thermo min |E/RATE|) (Torch.cpp:304-339) -- chosen on the device, and the
 * Strang alternation of single-stage hydro sub-steps (Torch::subStep) and
 * thermo sub-steps (as torchgpu_thermo_grid_substep).  Both handles must use
 * the same stream.  One host round trip at the start of the call (the
 * step count, so that Torch::stepCounter turns only with steps that ran),
 * none inside the loop; errors end the run and surface at the next
 * synchronizing call. */
int torchgpu_run_split(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                       const double* F_FUV, int64_t excluded, int nsteps, int surrogate);

/* A star on the grid (Parameters.Star: cell_position_x/y/z,
 * wind_radius_in_cells, photon_rate; src/main.cpp:192-204), at the centre
 * of its cell (snap_to_face_left_* off).  photon_rate in code units. */
typedef struct torchgpu_star {
    int cell[3];
    int wind_cell_radius;
    double photon_rate;
} torchgpu_star;
/* Star-on thermodynamics (Star.on = true) for the grid of ctx: from now on
 * the grid thermo calls (grid_timestep, grid_substep, run_split,
 * grid_inputs) compute COL_DEN by the causal column-density ray trace of
 * Thermodynamics::rayTrace (Thermodynamics.cpp:339-417, on the state's
 * fixed primitives) where no col_den array is passed, take F_FUV =
 * fluxFUV(0.5 photon_rate, r^2) (Thermodynamics.cpp:237-243) where no F_FUV
 * array is passed, and skip the CausalWind cells (the star's cell and those
 * within the wind radius, Fluid.cpp:41-54) as preTimeStepCalculations does.
 * The neighbour geometry is Grid::calculateNearestNeighbours
 * (Grid.cpp:922-951) and Radiation::cellPathLength (Radiation.cpp:102-117).
 * NULL switches the star off.  The star cell must lie inside the grid. */
int torchgpu_cooling_set_star(torchgpu_cooling* c, torchgpu_ctx* ctx, const torchgpu_star* star);
/* Thermodynamics::initialiseMinTempField (Thermodynamics.cpp:102-111): the
 * per-cell T_min [n] the grid calls take when they are passed no T_min
 * (with min_temp_initial_state; NULL restores the constant m_T_min = 100).
 * Host array unless `on_device`.  Synchronises. */
int torchgpu_cooling_set_min_temp(torchgpu_cooling* c, const double* T_min, int64_t n, int on_device);
/* The star's per-cell geometry, as Radiation::initField sets it
 * (Radiation.cpp:78-96): the four upstream neighbours (flat indices, -1
 * none) and weights (int64 / double [n][4]), the path length ds and the
 * shell volume ([n]); host arrays unless `on_device`.  Synchronises. */
int torchgpu_star_geometry(torchgpu_cooling* c, int64_t* nb, double* w, double* ds, double* shell_vol, int on_device);
/* The per-cell inputs [9][n] (den, pre, hii, adv (-1: skipped), E, T_min,
 * COL_DEN, F_FUV, dt = 0) the grid thermo operator forms for the current
 * state, including the ray-traced COL_DEN and, into dcol_den [n] when not
 * NULL, its DCOL_DEN (Thermodynamics::preTimeStepCalculations' inputs).
 * Synchronises. */
int torchgpu_thermo_grid_inputs(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* T_min, const double* col_den,
                                const double* F_FUV, int64_t excluded, double* in9, double* dcol_den, int on_device);

/* GridCell::H, the heating arrays Torch::run prints at every checkpoint
 * (Torch.cpp:264,272-276): HID order IMLC, NMLC, RHII, CEHI, CIEC, NMC, EUVH,
 * FUVH, IRH, CRH, TOT (src/Torch/Common.hpp:32-34). */
enum { TORCHGPU_HEAT_IMLC = 0, TORCHGPU_HEAT_NMLC, TORCHGPU_HEAT_RHII, TORCHGPU_HEAT_CEHI, TORCHGPU_HEAT_CIEC,
       TORCHGPU_HEAT_NMC, TORCHGPU_HEAT_EUVH, TORCHGPU_HEAT_FUVH, TORCHGPU_HEAT_IRH, TORCHGPU_HEAT_CRH,
       TORCHGPU_HEAT_TOT, TORCHGPU_HEAT_N };
/* With `on`, the grid keeps the heating terms the reference keeps in
 * GridCell::H between fills (zero at set_primitive, as a new GridCell):
 * every thermo sub-step's Thermodynamics::integrate sets TOT to the
 * post-subcycle RATE of each non-wind cell or zeroes H below the HII switch
 * (Thermodynamics.cpp:275-280,334-335); every Radiation pre-step sets RHII
 * and EUVH (Radiation.cpp:278-279).  Off frees them. */
int torchgpu_heating_track(torchgpu_ctx* ctx, int on);
/* Thermodynamics::fillHeatingArrays (Thermodynamics.cpp:419-461) on the
 * state Torch::run prints (torchgpu_get_print_state): per non-wind cell the
 * eight heating and cooling terms from Q (cooling terms negated), RHII and
 * EUVH as kept, and TOT += RHII + EUVH (kept: it accumulates over fills until
 * the next integrate resets it); wind cells and cells below the HII switch
 * give zeros.  COL_DEN is the star's ray trace on Q's density when a star is
 * set and col_den is NULL (Thermodynamics.cpp:420-421), else col_den or zero;
 * F_FUV the star's flux, else F_FUV or zero.  Needs heating_track on before
 * the run.  H [11][n] code units, host unless `on_device`.  Synchronises. */
int torchgpu_fill_heating(torchgpu_ctx* ctx, torchgpu_cooling* c, const double* col_den, const double* F_FUV,
                          int64_t excluded, double* H, int on_device);
/* DataPrinter::printHeating (src/IO/DataPrinter.cpp:325-354): time t (s),
 * nx, ny, nz, then per cell (x fastest) the nd cell-centre coordinates in
 * cell units (index + 0.5, as the reference prints GridCell::xc) and the 11
 * terms of H (host [11][nz][ny][nx], code units) in erg cm^-3 s^-1, tab
 * separated, "%.10e"; gzip when `gz`. */
int torchgpu_heating_write(const char* path, int gz, double t, const double* H, int nx, int ny, int nz,
                           double density_scale, double pressure_scale, double time_scale);

/* ------------------------------------------------------------------------ */
/* Radiation (photoionisation by the star; src/Integrators/Radiation.cpp)     */
/* ------------------------------------------------------------------------ */
/* Radiation.integration_scheme and Radiation.coupling (Radiation.cpp:40-52,
 * enum order of src/Torch/Common.hpp:38-39). */
enum { TORCHGPU_RT_IMPLICIT = 0, TORCHGPU_RT_IMPLICIT2 = 1, TORCHGPU_RT_EXPLICIT = 2 };
enum { TORCHGPU_RT_TTI = 0, TORCHGPU_RT_NEQ = 1, TORCHGPU_RT_OFF = 2 };
/* Radiation::initialise's state in code units (fill with
 * torchgpu_rad_params_init): K1..K4, cross-section, tau_0, temperatures,
 * scheme, coupling, constants, the Hummer recombination (rec) and
 * recombination-cooling (cool) linear splines (LinearSplineData, 31 knots). */
typedef struct torchgpu_rad_params {
    double K[4];
    double sigma, tau0, min_hii, T_hi, T_hii;
    int scheme, coupling, collisions_on, pad;
    double heating_amplification, mass_fraction_h;
    double hydrogen_mass, spec_gas_const, boltzmann, rydberg, pi;
    double voronov_A, voronov_X, voronov_K, voronov_P, voronov_DE;
    double photon_energy, T_min;
    double rec_x[31], rec_y[31], rec_y2[31], rec_slope[2];
    double cool_y[31], cool_y2[31], cool_slope[2];
} torchgpu_rad_params;
/* Radiation::initialise + initRecombinationHummer (Radiation.cpp:24-56,
 * 184-212) from the Lua-level values: Radiation.K1..K4,
 * photoion_cross_section (cm^2), tau_0, minimum_hii_fraction,
 * temperature_hi/hii (K), integration_scheme, coupling, collisions_on,
 * heating_amplification, mass_fraction_hydrogen and Star.photon_energy
 * (erg), with the unit scales.  T_min is the reference's default cell
 * minimum temperature, 100 K. */
int torchgpu_rad_params_init(torchgpu_rad_params* r, double density_scale, double pressure_scale, double time_scale,
                             const double K[4], double photoion_cross_section, double tau0, double min_hii_fraction,
                             double temperature_hi, double temperature_hii, int scheme, int coupling,
                             int collisions_on, double heating_amplification, double mass_fraction_h,
                             double photon_energy);

typedef struct torchgpu_radiation torchgpu_radiation;
/* The Radiation integrator on the grid of ctx with a star (Star.on; its
 * photon_rate in code units): per-cell state R [6][n] (HII_A, TAU, TAU_A,
 * DTAU, DTAU_A, HEAT, the reference's RID order) on the device, the star
 * geometry of torchgpu_star_geometry.  Uses ctx's stream; errors surface
 * at torchgpu_synchronize(ctx) as TORCHGPU_ERR_RADIATION (or the
 * fixSolution error). */
int torchgpu_radiation_create(const torchgpu_rad_params* p, torchgpu_ctx* ctx, const torchgpu_star* star,
                              torchgpu_radiation** out);
void torchgpu_radiation_destroy(torchgpu_radiation* r);
/* Radiation::initField (Radiation.cpp:78-96) as Torch::initialise calls it
 * after globalUfromQ (Torch.cpp:121): DTAU and DTAU_A from the density and
 * HII fraction of Q -- the primitives [7][n] given to set_primitive, fixed
 * as there (host unless `on_device`), or with Q NULL those of
 * fixPrimitives(QfromU(U)); the rest of R zero. */
int torchgpu_radiation_init_field(torchgpu_radiation* r, const double* Q, int on_device);
/* Q = fixPrimitives(QfromU(U)) and Radiation::preTimeStepCalculations
 * (heating, Radiation.cpp:257-296) on the current state, as Torch::fullStep
 * does before the time step (Torch.cpp:390-395).  Asynchronous. */
int torchgpu_radiation_pre_step(torchgpu_radiation* r);
/* Radiation::calculateTimeStep (Radiation.cpp:298-351; the reference's
 * first-call rule for K1 included, so call it only where Torch::
 * calculateTimeStep does -- not on the first step) on the state of the last
 * pre_step (made now if there is none) -> *dt (<= dt_max).  Synchronises. */
int torchgpu_radiation_timestep(torchgpu_radiation* r, double dt_max, double* dt);
/* Torch::subStep of the radiation component (Torch.cpp:352-364): checkValues,
 * unless `has_heat` (the state is the one of the last timestep call) Q and
 * preTimeStepCalculations again, integrate (transferRadiation for the
 * implicit scheme, Radiation.cpp:569-651, else transferRadiation2,
 * :510-567: the causal sweep from the star), updateSourceTerms (:653-680),
 * advSolution(dt), fixSolution.  Asynchronous. */
int torchgpu_radiation_substep(torchgpu_radiation* r, double dt, int has_heat);
/* R [6][n] and Q[HII], Q[ADV] [2][n] as the last sub-step left them; host
 * arrays unless `on_device`.  Synchronises. */
int torchgpu_radiation_get_state(torchgpu_radiation* r, double* R, double* q_hii_adv, int on_device);
/* torchgpu_run_split with any of the reference's split components: the
 * cooling handle `c` and/or the radiation handle `r` (either may be NULL;
 * both NULL is torchgpu_run).  Each step: the preTimeStepCalculations of
 * thermodynamics and radiation, dt = min(dt_nextCheckpoint, dt_max * 1e-20
 * first, then min(min(dt_hydro, dt_rad), dt_thermo)) chosen on the device
 * (Torch.cpp:304-339,388-397), then the Strang sub-steps over (hydro,
 * thermo, radiation) rotated by stepCounter (Torch.cpp:399-415), each a
 * Torch::subStep.  With radiation, the state printed at checkpoints carries
 * the HII fraction its last sub-step left in Fluid::Q.  All handles must
 * use the same stream.  Asynchronous, as torchgpu_run_split. */
int torchgpu_run_full(torchgpu_ctx* ctx, torchgpu_cooling* c, torchgpu_radiation* r, const double* T_min,
                      const double* col_den, const double* F_FUV, int64_t excluded, int nsteps, int surrogate);

/* ------------------------------------------------------------------------ */
/* Cooling surrogate (no reference counterpart: a learned replacement of the  */
/* subcycled integrate, with the FP64 integrator as gate and fallback)       */
/* ------------------------------------------------------------------------ */

/* MLP 8 -> 128 -> 128 -> 128 -> 128 -> 3 (SiLU), bf16 weights, fp32
 * accumulation.  Parameter blobs (fp32, host): W = W1[128][8] | W2 | W3 |
 * W4 [128][128] | W5[3][128]; b = b1 | b2 | b3 | b4 [128] | b5[3]. */
enum { TORCHGPU_MLP_NW = 128 * 8 + 3 * 128 * 128 + 3 * 128, TORCHGPU_MLP_NB = 4 * 128 + 3 };
typedef struct torchgpu_mlp torchgpu_mlp;
int torchgpu_mlp_create(int max_batch, void* stream, torchgpu_mlp** out);
void torchgpu_mlp_destroy(torchgpu_mlp* m);
int torchgpu_mlp_set_params(torchgpu_mlp* m, const float* W, const float* b);
int torchgpu_mlp_get_params(torchgpu_mlp* m, float* W, float* b);
/* y[B][3] = MLP(feat[B][8]) with the fused inference kernel (device ptrs).
 * Asynchronous on the handle's stream: feat must be ready on that stream and
 * y is valid once it has been synchronized (pass the caller's stream at
 * create time to order the call with the caller's own work). */
int torchgpu_mlp_forward(torchgpu_mlp* m, const float* feat, int64_t B, float* y);
/* Loss 0.5/B sum_o w_o (y_o - t_o)^2 (output 2's target is the log10
 * relative error of output 0 when err_head != 0) and its gradients, without
 * an update.  feat [B][8], target [B][3] device; gW/gb host (may be NULL). */
int torchgpu_mlp_loss_grad(torchgpu_mlp* m, const float* feat, const float* target, int B, const float* w3,
                           int err_head, float* loss, float* gW, float* gb);
/* One Adam step on a batch (loss_grad + update of the master weights). */
int torchgpu_mlp_train_step(torchgpu_mlp* m, const float* feat, const float* target, int B, const float* w3,
                            int err_head, float lr, float weight_decay, float* loss);

/* Surrogate configuration of a cooling operator: network, feature
 * normalisation f = (u - mu) * isd of the raw features u = (log10 rho,
 * log10 T, x, log10 max(F_FUV, 1), Av, log10 dt, log10(subcycle count + 1),
 * log10(Tc / T) of the coarse end-temperature estimate Tc) in code units
 * (the last two from the FP64 prepass), the training envelope [lo, hi] of
 * the first seven (a cell outside it goes to the FP64 fallback; +-HUGE_VAL
 * leaves a feature unbounded), the error gate on the estimated relative T
 * error and the subcycle count up to which cells are integrated exactly. */
typedef struct torchgpu_surrogate_cfg {
    float mu[8], isd[8];
    double lo[7], hi[7];
    double gate;
    int nexact;
    int pad;
} torchgpu_surrogate_cfg;
int torchgpu_cooling_set_surrogate(torchgpu_cooling* c, torchgpu_mlp* m, const torchgpu_surrogate_cfg* cfg);
/* The surrogate step: out as torchgpu_thermo_cells; path (device, uint8
 * [n], may be NULL): 0 exact FP64, 1 surrogate, 2 FP64 fallback.  counts
 * (host, may be NULL, synchronizes): cells on the surrogate and fallback
 * paths. */
int torchgpu_thermo_cells_surrogate(torchgpu_cooling* c, int64_t n, const double* in, double* out, uint8_t* path,
                                    int64_t* counts);
/* Subcycle count of each cell (count, device int64 [n]) that
 * torchgpu_thermo_cells would integrate: the explicit steps after the first
 * (Thermodynamics::integrate, Thermodynamics.cpp:303-312, fixed by the
 * pre-step rate), 0 for a cell that is not subcycled, -1 where dt / dti is
 * beyond the int range; nothing is integrated and no error is raised (no
 * reference counterpart: it prices a call, or filters training samples,
 * before the integration).  Asynchronous. */
int torchgpu_thermo_subcycles(torchgpu_cooling* c, int64_t n, const double* in, int64_t* count);
/* Training data: for the cells with more than nexact subcycles, writes their
 * indices (list, device int64 [n]), features (feat [n][8], normalised with
 * the configured mu/isd), raw features (raw [n][8], may be NULL) and, from
 * the FP64 integrator's output `out`, the targets (target [n][3]:
 * d log10 T_sub, 0, 0).  Returns the count. */
int torchgpu_surrogate_training_data(torchgpu_cooling* c, int64_t n, const double* in, const double* out,
                                     int64_t* list, float* feat, float* target, float* raw, int64_t* count);

#ifdef __cplusplus
}
#endif
#endif
