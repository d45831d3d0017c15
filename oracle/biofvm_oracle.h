/*
 * TEST INFRASTRUCTURE ONLY — the CPU oracle.
 *
 * A plain-C restatement of paraBioFVM's host (OpenMP) solver path, used by the
 * test-suite as the parity checker and by bench.py as the timed CPU baseline.
 * No product code links, loads or calls it.  The same source is compiled twice:
 * with REAL = float  (liboracle_f32.so, the reference's default build) and
 * with REAL = double (liboracle_f64.so, the reference's -DDOUBLE_PRECISION build,
 * CMakeLists.txt:64-66), always with -O3 -fopenmp -ffp-contract=off.
 *
 * Every function cites the reference file:line it restates.  Layouts follow
 * src/traits.h:65-93: densities [z][y][x][s]; 3-D gradients [d][z][y][x][s];
 * 1-D/2-D gradients [y][x][s][d]; ballots [z][y][x].
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - diffusion, Dirichlet, gradient, secretion: the reference test-suite's
 *     known-answer and closed-form values (test/test.cpp), committed as data in
 *     tests/golden/reference_expectations.json;
 *   - Thomas sweeps: a dense float64 tridiagonal solve of the same LOD system;
 *   - voxel index: exhaustive boundary checks against float32 (p-min)/dx;
 *   - secretion with conflicts: a brute-force per-voxel reduction
 *     (test/utils.cpp:134-202 semantics);
 *   - release at death (or_cell_release): its closed form in FP64, D + I*f/vv at
 *     the agent's voxel and I cleared, a no-op with internalization off and for
 *     agents outside the mesh (the reference's suites never test release);
 *   - bulk supply/uptake (or_bulk_solve): its geometric closed forms in FP64.
 * Every exported compute function has one of these pins.
 */
#ifndef BIOFVM_ORACLE_H
#define BIOFVM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#ifdef ORACLE_DOUBLE
typedef double or_real;
#define OR_FN(name) name##_f64
#else
typedef float or_real;
#define OR_FN(name) name##_f32
#endif

typedef struct or_state or_state;

/* mesh.cpp:7-25 (grid shape) and microenvironment.cpp:27-53 (allocation, initial fill). */
or_state* OR_FN(or_create)(int dims, const int* mins, const int* maxs, const int* voxel_shape, int substrates,
						   or_real time_step, const or_real* diffusion_coefficients, const or_real* decay_rates,
						   const or_real* initial_conditions);
void OR_FN(or_destroy)(or_state* st);

void OR_FN(or_grid_shape)(const or_state* st, int* out3);
or_real* OR_FN(or_densities)(or_state* st);
or_real* OR_FN(or_gradients)(or_state* st);

/* microenvironment.h:47-50 boundary arrays; NULL values clears the face. */
void OR_FN(or_set_boundary)(or_state* st, int dim, int is_max, const or_real* values, const unsigned char* conds);
/* microenvironment.h:42-45 interior list (copied). */
void OR_FN(or_set_interior)(or_state* st, int count, const int* voxels, const or_real* values,
							const unsigned char* conds);
/* agent_data.h:11-34: SoA arrays are borrowed; internalized is updated in place. */
void OR_FN(or_set_agents)(or_state* st, int count, const or_real* secretion_rates, const or_real* saturation_densities,
						  const or_real* uptake_rates, const or_real* net_export_rates, or_real* internalized,
						  const or_real* fraction_released_at_death, const or_real* volumes, const or_real* positions);
void OR_FN(or_set_compute_internalized)(or_state* st, int flag);
void OR_FN(or_set_threads)(int threads);

/* host/solver.cpp:5-10: cell.initialize then diffusion.initialize. */
int OR_FN(or_initialize)(or_state* st);

/* host/diffusion_solver.cpp:72-142 — exposes the precomputed Thomas coefficients of one axis. */
int OR_FN(or_diffusion_coefficients)(const or_state* st, int axis, or_real* b_out, or_real* c_out, or_real* e_out);

int OR_FN(or_diffusion_solve)(or_state* st);		   /* host/diffusion_solver.cpp:60-70, 482-544 */
int OR_FN(or_dirichlet_solve)(or_state* st);		   /* host/dirichlet_solver.cpp:112-144 */
int OR_FN(or_gradient_solve)(or_state* st);		   /* host/gradient_solver.cpp:275-291 */
int OR_FN(or_cell_simulate)(or_state* st, int recompute); /* host/cell_solver.cpp:272-299 */
/* host/bulk_solver.cpp:23-35: supply, uptake, target are the callbacks' values as [V][S] arrays */
int OR_FN(or_bulk_solve)(or_state* st, const or_real* supply, const or_real* uptake, const or_real* target);
int OR_FN(or_cell_release)(or_state* st, int index);	   /* host/cell_solver.cpp:340-346 */

/* One timestep in a single omp parallel region (main.cpp:64-107 pattern), in the order
 * secretion -> diffusion -> gradient; phase_ms receives per-phase wall times. */
int OR_FN(or_timestep)(or_state* st, int recompute, int with_gradient, double* phase_ms);

/* mesh.cpp:47-53 for a batch of positions (dims floats each); out3 receives 3 ints per point. */
void OR_FN(or_voxel_positions)(const or_state* st, int count, const or_real* positions, int* out3);

const char* OR_FN(or_last_error)(void);

#ifdef __cplusplus
}
#endif

#endif
