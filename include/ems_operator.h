/*
 * ems_operator.h — host-side EC-FDTD operator builder (C ABI).
 *
 * Builds, from a CSXCAD-free scene of axis-aligned boxes and wires, exactly the arrays
 * the reference Operator hands to its engines:
 *   - the float32 update coefficients vv, vi, ii, iv in the reference's
 *     N-I-J-K layout (component, x, y, z; z fastest)
 *     [Operator::vv_ptr/vi_ptr/ii_ptr/iv_ptr, FDTD/operator.h:362-367]
 *   - the timestep [Operator::CalcTimestep, FDTD/operator.cpp:1882-2030]
 *   - the excitation signal and cell lists
 *     [Excitation, FDTD/excitation.cpp:95-276;
 *      Operator_Ext_Excitation::BuildExtension, operator_ext_excitation.cpp:105-297]
 *   - UPML slabs and coefficients [Operator_Ext_UPML, operator_ext_upml.cpp:69-445]
 *   - Mur faces and coefficients [Operator_Ext_Mur_ABC, operator_ext_mur_abc.cpp:80-186]
 *   - lumped parallel-RC elements folded into the operator
 *     [Operator::Calc_LumpedElements, FDTD/operator.cpp:1586-1763]
 *
 * Replaces Operator::SetGeometryCSX + Operator::CalcECOperator
 * (FDTD/operator.cpp:822-837, 986-1097) for scenes of boxes, cylinders,
 * material grids and wires (polyline curves).
 * All functions return 0 on success and a negative EMSOP_E_* code on error;
 * emsop_last_error() describes the last failure of the calling thread.
 */
#ifndef EMS_OPERATOR_H
#define EMS_OPERATOR_H

#ifdef __cplusplus
extern "C" {
#endif

#define EMSOP_OK 0
#define EMSOP_E_ARG (-1)
#define EMSOP_E_STATE (-2)
#define EMSOP_E_NUMERIC (-3)
#define EMSOP_E_NOMEM (-4)

typedef struct emsop emsop;

/* signal types, Excitation::ExciteTypes (FDTD/excitation.h:29) */
#define EMSOP_SIG_GAUSS 0
#define EMSOP_SIG_SINUS 1
#define EMSOP_SIG_DIRAC 2
#define EMSOP_SIG_STEP 3
#define EMSOP_SIG_CUSTOM 10

const char* emsop_last_error(void);

int emsop_create(emsop** out);
void emsop_destroy(emsop* h);

/* mesh lines in drawing units and the drawing unit in meters
 * (CSRectGrid lines + delta unit, Operator::SetupCSXGrid FDTD/operator.cpp:793-820) */
int emsop_set_grid(emsop* h, const double* x, unsigned nx, const double* y, unsigned ny,
                   const double* z, unsigned nz, double delta_unit);
int emsop_set_background(emsop* h, double eps_r, double mue_r, double kappa, double sigma);
/* material box: per-direction eps_r, mue_r, kappa (S/m), sigma (Ohm/m) */
int emsop_add_material(emsop* h, const double eps[3], const double mue[3], const double kappa[3],
                       const double sigma[3], int priority, const double start[3], const double stop[3]);
/* material grid: ncells[0] x ncells[1] x ncells[2] mesh cells from cell
 * cell0 (cell i spans lines i..i+1) with eps_r and kappa (S/m, may be NULL)
 * per cell, x fastest, isotropic; mue_r, sigma uniform.  Equivalent to
 * one-cell material boxes of this priority added in x-fastest order (a point
 * on a face shared by grid cells takes the lowest cell index); replaces them
 * for large design regions.  The grid must be set first. */
int emsop_add_material_grid(emsop* h, const unsigned cell0[3], const unsigned ncells[3], const double* eps_r,
                            const double* kappa, double mue_r, double sigma, int priority);
int emsop_add_metal(emsop* h, int priority, const double start[3], const double stop[3]);
/* Lorentz / Drude dispersion of the material box added last (CSPropLorentzMaterial,
 * built by Operator_Ext_LorentzMaterial::BuildExtension,
 * FDTD/extensions/operator_ext_lorentzmaterial.cpp:120-445): params[(o*6 + k)*3 + n]
 * for order o < norder, direction n, k = eps plasma frequency (Hz), eps
 * relaxation time (s), eps Lorentz pole frequency (Hz), mue plasma
 * frequency, mue relaxation time, mue Lorentz pole frequency; 0 = unset. */
int emsop_set_dispersion(emsop* h, unsigned norder, const double* params);
/* Debye dispersion of the material box added last (CSPropDebyeMaterial, same
 * builder): params[(o*2 + k)*3 + n], k = eps delta, relaxation time (s);
 * a relaxation time below 2 dT leaves the order off for that edge. */
int emsop_set_debye(emsop* h, unsigned norder, const double* params);
/* After emsop_build: for the edge (component n, pos), the four cells of its
 * quarter-cell average (cells[3*q + axis], 0xFFFFFFFF for a cell before the
 * first line; q = up-right, up-left, down-right, down-left in the
 * (n+1, n+2) plane) and dC/d eps_r of each, where
 * C = eps0 * sum_q eps_q A_q / sum_q A_q * EdgeArea / EdgeLength
 * (Operator::AverageMatQuarterCell + Calc_ECPos, FDTD/operator.cpp:1189-1254,1447-1510);
 * dG/d kappa = dC/d eps_r / eps0.  Quarter-cell averaging only. */
int emsop_eps_weights(const emsop* h, int n, const unsigned pos[3], unsigned cells[12], double dC[4]);
/* lumped element; parallel=1 for the R||C form folded into the operator */
int emsop_add_lumped(emsop* h, int dir, double R, double C, double L, int caps, int parallel,
                     int priority, const double start[3], const double stop[3]);
/* excitation box: type 0 soft E, 1 hard E, 2 soft H, 3 hard H; val = per-direction amplitude */
int emsop_add_excitation(emsop* h, int type, const double val[3], double delay, int priority,
                         const double start[3], const double stop[3]);
/* Polyline wire of a metal property (CSPrimCurve): npts points pts[3*i + axis]
 * in drawing units; every segment's mesh path (emsop_find_path) gets
 * vv = vi = 0 (Operator::CalcPEC_Curves, FDTD/operator.cpp:2086-2118). */
int emsop_add_metal_curve(emsop* h, unsigned npts, const double* pts);
/* Polyline excitation (CSPrimCurve of an excitation property): type 0 soft or
 * 1 hard voltage excitation on the path edges of every segment, amplitude
 * val[n] times the edge length, entered after all box excitations in the
 * order added (Operator_Ext_Excitation::BuildExtension,
 * FDTD/extensions/operator_ext_excitation.cpp:235-287); types 2 and 3 add nothing. */
int emsop_add_excitation_curve(emsop* h, int type, const double val[3], double delay, unsigned npts,
                               const double* pts);
/* Cylinder of a metal property (CSPrimCylinder, shell_width = 0) or
 * cylindrical shell (CSPrimCylindricalShell: radius is the wall's mid radius,
 * shell_width its thickness), axis start -> stop in drawing units.  A point is
 * inside when the foot of its perpendicular on the axis lies within the axis
 * segment and its distance d from the axis satisfies d <= radius (cylinder)
 * or |d - radius| <= shell_width / 2 (shell).  It takes part in the
 * MATERIAL|METAL lookup at every edge centre like a metal box
 * (Operator::CalcPEC_Range, FDTD/operator.cpp:2045-2083). */
int emsop_add_metal_cylinder(emsop* h, int priority, const double start[3], const double stop[3], double radius,
                             double shell_width);
/* Cylinder or cylindrical shell of an excitation property, as
 * emsop_add_excitation with that shape (operator_ext_excitation.cpp:150-215). */
int emsop_add_excitation_cylinder(emsop* h, int type, const double val[3], double delay, int priority,
                                  const double start[3], const double stop[3], double radius, double shell_width);
/* Weight functions of the excitation primitive added last (box, cylinder or
 * curve; CSPropExcitation::SetWeightFunction): weight[n] is an expression in
 * x, y, z (drawing units), rho = sqrt(x^2 + y^2), r = sqrt(x^2 + y^2 + z^2)
 * and a = atan2(y, x), in the syntax of emsop_set_custom_signal, or NULL for
 * weight 1.  The amplitude of an edge of direction n becomes
 * (weight[n](Yee point) * val[n]) * edge length
 * (operator_ext_excitation.cpp:176,212,267). */
int emsop_set_excitation_weight(emsop* h, const char* const weight[3]);
/* Evaluates an expression of the syntax above at npts points: vars names the
 * nvar variables, comma separated; values[k * nvar + i] is variable i at point
 * k; out[k] the result.  Used for the mode functions of mode-matching probes
 * (ProcessModeMatch::InitProcess, Common/processmodematch.cpp:78-147). */
int emsop_eval_function(const char* text, const char* vars, unsigned nvar, const double* values, unsigned npts,
                        double* out);
/* bc: 0 PEC, 1 PMC, 2 Mur, 3 PML, -1 none (openEMS::m_BC_type) */
int emsop_set_boundary(emsop* h, const int bc[6], const unsigned pml_size[6], const double mur_vph[6]);
int emsop_set_signal(emsop* h, int type, double f0, double fc);
/* custom excitation (Excitation::SetupCustomExcite / CalcCustomExcitation,
 * FDTD/excitation.cpp:86-93,219-252): NrTS samples of the expression f(t) at
 * t = n dT (voltages) and (n + 1/2) dT (currents), Nyquist number from f0.
 * Syntax of the reference's function parser: numbers, t, pi, e, + - * / ^,
 * comparisons, & | !, sin cos tan asin acos atan sinh cosh tanh exp log log10
 * sqrt abs floor ceil int, atan2 pow min max, if(c, a, b).  A parse error
 * fails with EMSOP_E_ARG and its position. */
int emsop_set_custom_signal(emsop* h, const char* function, double f0);
/* method: 1 or 3 (Rennings); forced_dt <= 0 means automatic; factor in (0,1] */
int emsop_set_timestep(emsop* h, int method, double forced_dt, double factor);
int emsop_set_cell_constant_material(emsop* h, int enable);
/* threads used by the build (0 = all hardware threads) */
int emsop_set_threads(emsop* h, int nthreads);

/* run the operator build for a simulation of nrts timesteps */
int emsop_build(emsop* h, unsigned nrts);

/* ---- results (valid after emsop_build) ---- */
int emsop_num_lines(const emsop* h, unsigned out[3]);
double emsop_timestep(const emsop* h);
unsigned emsop_nyquist(const emsop* h);
double emsop_signal_period(const emsop* h);
unsigned emsop_signal_length(const emsop* h);
/* period in timesteps as Engine_Ext_Excitation computes it, 0 when aperiodic */
int emsop_signal_period_steps(const emsop* h);
const float* emsop_signal_volt(const emsop* h);
const float* emsop_signal_curr(const emsop* h);
/* which: 0 vv, 1 vi, 2 ii, 3 iv; [3][Nx][Ny][Nz] float32 */
const float* emsop_coeff(const emsop* h, int which);
/* which: 0 C, 1 G, 2 L, 3 R equivalent-circuit arrays (float32, N-I-J-K) */
const float* emsop_ec(const emsop* h, int which);
/* effective material of edge (ny, pos): out = eps*EPS0, kappa, mue*MUE0, sigma
 * (Operator::Calc_EffMatPos, FDTD/operator.cpp:1566-1584) */
int emsop_effmat_at(const emsop* h, int ny, const unsigned pos[3], double out[4]);

unsigned emsop_exc_count(const emsop* h, int curr);
/* idx: [3][K] uint32, dir: [K] uint16, amp [K] float32, delay [K] uint32 */
int emsop_exc_get(const emsop* h, int curr, unsigned* idx, unsigned short* dir, float* amp, unsigned* delay);

unsigned emsop_upml_count(const emsop* h);
/* arrays: vv, vvfo, vvfn, ii, iifo, iifn, each [3][num0][num1][num2] */
int emsop_upml_get(const emsop* h, unsigned k, unsigned start[3], unsigned num[3], const float* arrays[6]);

/* Operator_Ext_LorentzMaterial tables (FDTD/extensions/operator_ext_lorentzmaterial.cpp:120-445):
 * number of dispersion orders, and per order the cell count, the flags
 * on = {volt ADE, curr ADE, volt Lorentz pole, curr Lorentz pole}, the cell
 * positions pos[3][count] and tabs[3*k + n] = v_int, v_ext, i_int, i_ext,
 * v_lor, i_lor of direction n ([count] float32, valid only where the
 * matching flag is on).  Pointers stay valid until emsop_destroy. */
unsigned emsop_lorentz_orders(const emsop* h);
int emsop_lorentz_get(const emsop* h, unsigned order, unsigned* count, int on[4], const unsigned* pos[3],
                      const float* tabs[18]);
/* Operator_Ext_LumpedRLC tables (FDTD/extensions/operator_ext_lumpedRLC.cpp:112-548):
 * element edges of parallel RLC (L > 0) and series RLC elements, pos[3][K],
 * dir[K], coef[9][K] = i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2 (float32).
 * Their edges' EC values and coefficients are rewritten in the operator. */
unsigned emsop_rlc_count(const emsop* h);
int emsop_rlc_get(const emsop* h, const unsigned* pos[3], const int** dir, const float* coef[9]);

unsigned emsop_mur_count(const emsop* h);
/* coefficients [N[nyP]][N[nyPP]] for the two tangential components */
int emsop_mur_get(const emsop* h, unsigned k, int* ny, unsigned* line, unsigned* line_shift,
                  unsigned* start_ts, const float** coeff_nyP, const float** coeff_nyPP);

/* mesh queries used by probe setup */
double emsop_disc_line(const emsop* h, int n, unsigned pos, int dual);
double emsop_edge_length(const emsop* h, int n, unsigned pos, int dual);
/* SnapToMesh; returns 1 if inside on all axes */
int emsop_snap_to_mesh(const emsop* h, const double coord[3], int dual, unsigned out[3], int inside[3]);
/* SnapBox2Mesh (method 0): returns dimension or -2 outside */
int emsop_snap_box(const emsop* h, const double start[3], const double stop[3], int dual,
                   unsigned ustart[3], unsigned ustop[3], int start_in[3], int stop_in[3]);
/* Operator::FindPath (FDTD/operator.cpp:390-496): the staircase of mesh edges
 * between the nodes nearest start and stop; edge t is component dir[t] at node
 * pos[3*t + axis].  *count is set to the number of edges (0: no path); fails
 * with EMSOP_E_ARG if it exceeds max_edges.  pos and dir may be NULL. */
int emsop_find_path(const emsop* h, const double start[3], const double stop[3], unsigned max_edges,
                    unsigned* pos, int* dir, unsigned* count);

/* release the EC and effective-material arrays once extensions are built */
void emsop_release_build_arrays(emsop* h);

#ifdef __cplusplus
}
#endif
#endif /* EMS_OPERATOR_H */
