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
 *   GridCell::U state read/write (src/Fluid/GridCell.hpp:265)  torchgpu_set_conserved /
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
 *   Thermodynamics operator on the grid state             torchgpu_thermo_substep
 *   MLP surrogate of the subcycled update (no reference   torchgpu_mlp_*
 *     counterpart; FP64 integrate remains the fallback)
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
    TORCHGPU_ERR_UNKNOWN_LIMITER = 8,/* SlopeLimiterFactory::create unknown type (SlopeLimiter.cpp:331) */
    TORCHGPU_ERR_NSTEPS_RANGE = 9,   /* subcycle count beyond int range (Thermodynamics.cpp:307) */
    TORCHGPU_ERR_NOT_TRAINED = 10    /* surrogate used before weights were set */
};

/* Boundary conditions, Condition enum order of src/Torch/Common.hpp:27. */
enum { TORCHGPU_BC_FREE = 0, TORCHGPU_BC_REFLECTING = 1, TORCHGPU_BC_OUTFLOW = 2,
       TORCHGPU_BC_INFLOW = 3, TORCHGPU_BC_PERIODIC = 4 };
/* Riemann solvers, RiemannSolverFactory::create strings "HLL", "HLLC",
 * "RotatedHLLC" ("default"), src/Integrators/Riemann.cpp:246-257. */
enum { TORCHGPU_SOLVER_HLL = 0, TORCHGPU_SOLVER_HLLC = 1, TORCHGPU_SOLVER_ROTATED_HLLC = 2 };
/* Slope limiters, SlopeLimiterFactory::create strings
 * "monotonised_central", "superbee", "minmod", "maxmod", "leer", "ospre",
 * "albada" ("default"), src/Integrators/SlopeLimiter.cpp:313-332. */
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
 * the current HIP device.  `stream` is a hipStream_t (NULL: the library
 * creates one). */
int torchgpu_create(const torchgpu_params* params, void* stream, torchgpu_ctx** out);
void torchgpu_destroy(torchgpu_ctx* ctx);
int torchgpu_last_error(const torchgpu_ctx* ctx, char* buf, size_t len);
/* Device address of the conserved state U ([7][nz][ny][nx] float64). */
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
/* Up to `nsteps` iterations of the Torch::run loop for a hydro-only run,
 * with dt chosen on the device (first step dt_max * 1e-20, then the CFL
 * min-reduction fused into the corrector stage, clipped to the next
 * checkpoint).  Asynchronous. */
int torchgpu_run(torchgpu_ctx* ctx, int nsteps);
/* Restart the Torch::run bookkeeping (time 0, first-step rule, checkpoints). */
int torchgpu_reset_clock(torchgpu_ctx* ctx);
/* Waits for the stream and reports the first device-side error. */
int torchgpu_synchronize(torchgpu_ctx* ctx);
/* Simulation time, dt of the last step and number of completed steps. */
int torchgpu_get_clock(torchgpu_ctx* ctx, double* time, double* last_dt, int* steps);

#ifdef __cplusplus
}
#endif
#endif
