/*
 * torch_oracle.h -- CPU FP64 restatement of TORCH's hydro + thermodynamics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle that the
 * HIP path is checked against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in torch_amd/ links or
 * calls it; the product path never falls back to it.
 *
 * Every function follows the reference (hgsteggles/TORCH, src/...) arithmetic
 * operation by operation, in the same order, so that results are bitwise
 * equal to the reference's own FP64 build (g++ -O3, x86-64, no FMA
 * contraction).  Citations are given per function in torch_oracle.c.
 *
 * Parity pinning (what checks each part):
 *   - hydro (limiter, UfromQ/QfromU/FfromQ, HLL, HLLC, rotated solver, BCs,
 *     fixPrimitives/fixSolution, CFL, predictor-corrector, driver): bitwise
 *     against fixtures written by the reference's own translation units
 *     (oracle/_ref harness) in tests/golden/, plus closed forms and
 *     invariants in tests/test_oracle_*.py.
 *   - thermodynamics (units, rate functions, spline, preTimeStep, subcycle):
 *     against reference-harness fixtures (bitwise where only + - * / sqrt are
 *     involved; libm exp/log/pow calls are glibc's on both sides).
 *   - MLP reference (forward/backward): finite differences + torch autograd.
 *
 * Layout: structure-of-arrays, U[v*ncell + i + nx*(j + ny*k)], variable order
 * DEN, PRE, HII, ADV, VEL0, VEL1, VEL2 (reference src/Torch/Common.hpp:23-25).
 */
#ifndef TORCH_ORACLE_H_
#define TORCH_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_DEN = 0, ORC_PRE = 1, ORC_HII = 2, ORC_ADV = 3, ORC_VEL = 4, ORC_NV = 7 };

/* Boundary conditions (reference src/Torch/Common.hpp:27 Condition order). */
enum { ORC_BC_FREE = 0, ORC_BC_REFLECTING = 1, ORC_BC_OUTFLOW = 2, ORC_BC_INFLOW = 3, ORC_BC_PERIODIC = 4 };
/* Riemann solvers (reference Riemann.cpp:246-257). */
enum { ORC_HLL = 0, ORC_HLLC = 1, ORC_ROTATED_HLLC = 2 };
/* Slope limiters (reference SlopeLimiter.cpp:56-75). */
enum { ORC_LIM_MC = 0, ORC_LIM_SUPERBEE = 1, ORC_LIM_MINMOD = 2, ORC_LIM_MAXMOD = 3,
       ORC_LIM_LEER = 4, ORC_LIM_OSPRE = 5, ORC_LIM_ALBADA = 6 };

typedef struct {
    int nx, ny, nz;
    int spatial_order;        /* 0 or 1 (1 = linear reconstruction)          */
    int solver;               /* ORC_HLL / ORC_HLLC / ORC_ROTATED_HLLC        */
    int limiter;              /* ORC_LIM_*                                    */
    int bc_left[3], bc_right[3];
    double dx;                /* cell width, code units                       */
    double gamma;
    double dfloor, pfloor, tfloor;
    double mass_frac_h;       /* X                                            */
    double spec_gas_const;    /* R in code units                              */
    double dt_max;
    int nthreads;             /* OpenMP threads for the cell loops (>=1)      */
} orc_hydro_params;

/* ---- per-face / per-cell primitives (exposed for unit tests) ---------- */
double orc_limiter(int limiter, double a, double b);
void orc_UfromQ(double u[7], const double q[7], double gamma);
void orc_QfromU(double q[7], const double u[7], double gamma);
void orc_FfromQ(double f[7], const double q[7], double gamma, int dim);
/* One face flux: solver(F, Ql, Qr, a_l2, a_r2, gamma, dim).  Returns 0, or 1
 * if the HII flux is NaN (the reference throws there). */
int orc_riemann(int solver, double F[7], const double Ql[7], const double Qr[7],
                double a_l2, double a_r2, double gamma, int dim);
void orc_riemann_batch(int solver, int n, const double* Ql, const double* Qr,
                       const int* dims, double gamma, double* F, int* status);

/* ---- whole-grid operations (interior cells only, SoA) ------------------ */
void orc_globalQfromU(const orc_hydro_params* p, const double* U, double* Q);
void orc_fixPrimitives(const orc_hydro_params* p, double* Q, const double* U);
/* Returns number of cells that would make the reference throw. */
int orc_fixSolution(const orc_hydro_params* p, double* U);
double orc_hydro_timestep(const orc_hydro_params* p, const double* Q);
/* UDOT = flux divergence of Q (BCs + reconstruction + Riemann + sources).
 * Returns nonzero if a Riemann solve produced a NaN HII flux. */
int orc_hydro_rhs(const orc_hydro_params* p, const double* Q, double* UDOT);

/* Torch::hydroStep(dt, true) on U (Q must be fixPrimitives(QfromU(U))).
 * W, Q2, UDOT: caller-provided work arrays of 7*ncell doubles. */
int orc_hydro_step(const orc_hydro_params* p, double* U, double* Q, double* W,
                   double* UDOT, double dt);

/* Torch::run for a hydro-only problem, capped at nsteps.  tmax is code time,
 * ncheckpoints as in the reference.  dts (nsteps) receives each dt.  With
 * debug (Integration.debug), Torch::calculateTimeStep's test ends the loop
 * after a step whose hydro dt has 100 dt / tmax <= 1e-6 (Torch.cpp:321-332);
 * *steps_out / *quit_out (may be NULL) receive the steps run and whether the
 * loop quit that way.
 * Returns 0 or an error code (1 NaN flux, 2 fixSolution failure, 3 checkValues). */
int orc_run_hydro(const orc_hydro_params* p, double* U, int nsteps, double tmax,
                  int ncheckpoints, double* dts, double* time_out, int debug, int* steps_out, int* quit_out);
/* The same from the start time t0 (the time of an initial-conditions
 * snapshot, Torch.cpp:58), with the snapshots Torch::run prints
 * (Torch.cpp:254,270-281,295-297): up to max_prints states Q (7n each) with
 * their checkpoint count and time; *nprint receives how many. */
int orc_run_hydro_print(const orc_hydro_params* p, double* U, int nsteps, double t0, double tmax,
                        int ncheckpoints, double* dts, double* time_out, double* qprint, int max_prints,
                        int* pcount, double* ptime, int* nprint, int debug, int* steps_out, int* quit_out);

/* ---- units (reference Converter.cpp / Constants.cpp) ------------------- */
typedef struct { double M, L, T; } orc_converter;
void orc_converter_set_rho_pressure_time(orc_converter* c, double rho, double pre, double t);
double orc_to_code_units(const orc_converter* c, double val, double mi, double li, double ti);
double orc_from_code_units(const orc_converter* c, double val, double mi, double li, double ti);

/* ---- thermodynamics (reference Thermodynamics.cpp) --------------------- */
typedef struct {
    /* constants in code units */
    double hydrogen_mass, spec_gas_const, boltzmann, dust_xsec, pi;
    double z0, T1, T2, T3, T4, imlc, nmlc, ciec_minT, ciec, cxhi_damp, n0, nmc;
    double fuvh_a, fuvh_b, fuvh_c, irh_a, irh_b, crh, T_min_default;
    /* log-log natural spline of the HI collisional-excitation rate */
    double spl_x[26], spl_y[26], spl_y2[26], spl_min_slope, spl_max_slope;
    /* switches */
    double thermo_hii_switch, heating_amplification, mass_frac_h, gamma, pfloor;
    int subcycling;
} orc_thermo_params;

void orc_thermo_init(orc_thermo_params* t, const orc_converter* c, double thermo_hii_switch,
                     double heating_amplification, double mass_frac_h, double gamma,
                     double pfloor, int subcycling);
double orc_log_spline(const orc_thermo_params* t, double x);
/* Individual rates (for unit tests); T in K, densities in code units. */
void orc_thermo_rates(const orc_thermo_params* t, double nH, double hii, double T,
                      double Av, double F_FUV, double out[9]);
/* Per-cell preTimeStepCalculations + integrate.  Inputs are the cell's Q
 * (den, pre, hii, adv), U[PRE] (total energy), T_min, COL_DEN and F_FUV.
 * Outputs: heat (T[HEAT]), rate_pre (T[RATE] after preTimeStep), rate
 * (T[RATE] after integrate), nsub (number of subcycle rate evaluations),
 * p_sub (subcycled pressure). */
void orc_thermo_cell(const orc_thermo_params* t, double den, double pre, double hii,
                     double adv, double E, double T_min, double col_den, double F_FUV,
                     double dt, double* heat, double* rate_pre, double* rate,
                     int64_t* nsub, double* p_sub);
void orc_thermo_batch(const orc_thermo_params* t, int64_t n, const double* in /* n x 9 */,
                      double* out /* n x 5 */, int nthreads);

/* Star geometry (Grid::calculateNearestNeighbours, Radiation::cellPathLength,
 * shellVolume) of one cell and of the whole grid (x fastest; nb, w 4 per
 * cell), the CausalWind cells, and the thermo column-density ray trace
 * (Thermodynamics::rayTrace); n = {nx, ny, nz}, star = its cell. */
void orc_star_cell_geometry(const int n[3], const int star[3], double dx, double pi, int i, int j, int k,
                            int64_t nb[4], double w[4], double* ds, double* shell_vol);
void orc_star_geometry(const int n[3], const int star[3], double dx, double pi, int64_t* nb, double* w, double* ds,
                       double* shell_vol);
void orc_star_wind(const int n[3], const int star[3], int wind_radius, unsigned char* wind);
void orc_star_fuv(const int n[3], const int star[3], double dx, double pi, double photon_rate, double* F);
/* Thermodynamics::fillHeatingArrays (Thermodynamics.cpp:419-461) on n cells:
 * Q [7][n] is Fluid::Q, col_den [n] T[COL_DEN] (NULL: 0), F [n] the F_FUV of
 * each cell (NULL: 0), wind [n] the CausalWind mask (cells left untouched);
 * heat [3][n] holds GridCell::H's TOT, RHII and EUVH across calls (in/out);
 * H [11][n] receives GridCell::H in HID order (zeros for wind cells). */
void orc_fill_heating(const orc_thermo_params* t, int64_t n, const double* Q, const double* col_den, const double* F,
                      const unsigned char* wind, double* heat, double* H);
void orc_thermo_ray_trace(const int n[3], const int star[3], const int64_t* nb, const double* w, const double* ds,
                          double hydrogen_mass, const double* den, double* col, double* dcol);

/* Radiation (src/Integrators/Radiation.cpp): parameters in code units,
 * the Hummer recombination splines, the per-cell state R [6][n] in the
 * reference's RID order. */
enum { ORC_R_HII_A = 0, ORC_R_TAU = 1, ORC_R_TAU_A = 2, ORC_R_DTAU = 3, ORC_R_DTAU_A = 4, ORC_R_HEAT = 5 };
enum { ORC_SCHEME_IMPLICIT = 0, ORC_SCHEME_IMPLICIT2 = 1, ORC_SCHEME_EXPLICIT = 2 };
enum { ORC_COUPLING_TTI = 0, ORC_COUPLING_NEQ = 1, ORC_COUPLING_OFF = 2 };
typedef struct {
    double K[4], sigma, alphaB, tau0, minX, THI, THII;
    int scheme, coupling, collisions_on;
    double heating_amplification, mass_frac_h;
    double hydrogen_mass, spec_gas_const, boltzmann, rydberg;
    double voronov_A, voronov_X, voronov_K, voronov_P, voronov_DE;
    double photon_energy, photon_rate;
    double rec_x[31], rec_y[31], rec_y2[31], rec_slope[2];
    double cool_x[31], cool_y[31], cool_y2[31], cool_slope[2];
} orc_rad_params;
void orc_rad_init(orc_rad_params* r, const orc_converter* c, const double K[4], double sigma, double alphaB,
                  double tau0, double minX, double THI, double THII, int scheme, int coupling, int collisions_on,
                  double heating_amplification, double mass_frac_h, double photon_energy, double photon_rate);
double orc_rad_recombination(const orc_rad_params* r, double T);
int64_t orc_first_nonwind_cell(const int n[3], const int star[3], int wind_radius);
void orc_rad_init_field(const orc_rad_params* r, const int n[3], const int star[3], double dx, double pi,
                        const double* Q, double* R);
int orc_rad_run(const orc_rad_params* r, const orc_hydro_params* p, const int star[3], int wind_radius,
                double T_min, int* first, double* U, double* R, double* Qout, int nsteps, double* dts);

#ifdef __cplusplus
}
#endif
#endif
