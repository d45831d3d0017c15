/*
 * parabiofvm_hip.h — C ABI of the MI355X (gfx950) solver library libparabiofvm_hip.so.
 *
 * One bfvm_ctx is the device side of one microenvironment: it owns every device
 * buffer and one HIP stream (default priority).  The entry points are what the
 * reference's device solver facade does per call, flattened to plain pointers
 * and sizes (no C++ or torch types):
 *
 *   reference (asmelko/paraBioFVM)                                 entry point
 *   -------------------------------------------------------------  ---------------------------
 *   device_context() ctor, two streams   CUDA/device_context.cpp:7-11    bfvm_create
 *   ~device_context                      CUDA/device_context.cpp:13-17   bfvm_destroy
 *   solver::initialize                   CUDA/solver.cpp:42-48,
 *     (device_context / cell / diffusion / dirichlet initialize:
 *      device_context.cpp:19-26, cell_solver.cpp:329-340,
 *      diffusion_solver.cpp:41-103, dirichlet_solver.cpp:17-88)         bfvm_initialize
 *   solver::store_data_to_solver         CUDA/solver.cpp:50,
 *     device_context::copy_to_device     device_context.cpp:56-97       bfvm_store_densities,
 *                                                                        bfvm_store_agents
 *   solver::load_data_from_solver        CUDA/solver.cpp:52,
 *     device_context::copy_to_host       device_context.cpp:99-116      bfvm_load_densities,
 *                                                                        bfvm_load_internalized,
 *                                                                        bfvm_load_gradients
 *   diffusion_solver::solve              host/diffusion_solver.cpp:60-70,
 *                                        CUDA/diffusion_solver.cpp:128-152 bfvm_diffusion_solve
 *   dirichlet_solver::solve              host/dirichlet_solver.cpp:112-120 bfvm_dirichlet_solve
 *   gradient_solver::solve (host-only in the reference device facade,
 *                                        host/gradient_solver.cpp:275-291) bfvm_gradient_solve
 *   cell_solver::simulate_secretion_and_uptake
 *                                        host/cell_solver.cpp:272-299,
 *                                        CUDA/cell_solver.cpp:249-286      bfvm_cell_simulate
 *   cell_solver::release_internalized_substrates
 *                                        host/cell_solver.cpp:340-346,
 *                                        CUDA/cell_solver.cpp:288-294      bfvm_cell_release
 *   solver::wait_for_all                 CUDA/solver.cpp:54-58             bfvm_wait_for_all
 *   one timestep of main_device.cpp:70-111 (secretion, diffusion, gradient)
 *                                        as a captured HIP graph            bfvm_timestep
 *
 * Every function returns 0 on success and a negative status on error; the
 * message of the last error on the calling thread is bfvm_last_error().  No
 * entry point aborts the process.  All enqueued work runs on the context's
 * stream in call order; bfvm_load_* and bfvm_wait_for_all synchronise it.
 *
 * Arrays follow the reference's layouts (src/traits.h:65-93): densities
 * [z][y][x][s]; gradients [d][z][y][x][s] for 3-D meshes and [y][x][s][d] for
 * 1-D/2-D meshes; agent arrays [agent][s]; positions [agent][dims]; the interior
 * Dirichlet list is [voxel][dims] int32 with values/conditions [voxel][s].
 * Boolean arrays are one byte per entry (C++ bool).
 */
#ifndef PARABIOFVM_HIP_H
#define PARABIOFVM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfvm_ctx bfvm_ctx;

/* Status codes. */
#define BFVM_OK 0
#define BFVM_ERR_ARGUMENT (-1)
#define BFVM_ERR_STATE (-2)
#define BFVM_ERR_HIP (-3)
#define BFVM_ERR_UNSUPPORTED (-4)

/* Density storage in HBM between kernels (bfvm_set_storage).  All arithmetic is FP32 in registers; values
 * that are set (uploads, Dirichlet) are rounded to nearest even, values that accumulate (sweep outputs,
 * secretion) stochastically.  BF16 stores the value itself; FP16 stores value * 2^k_s with a per-substrate
 * power of two chosen at each densities upload so that a bound on |density| of substrate s lands in
 * [2^10, 2^11): the largest of the uploaded densities, the Dirichlet values and the agents' saturation and
 * bulk target densities (the values the solvers move densities towards).  Net export and release at death
 * raise the bound as they are enqueued; when bound * 2^k_s reaches 2^14 the stored values and the scale are
 * divided by a power of two on the device before the next step, so FP16 never overflows. */
#define BFVM_STORAGE_F32 0
#define BFVM_STORAGE_BF16 1
#define BFVM_STORAGE_F16 2

/* Combining of agents that share a voxel (bfvm_set_secretion_reduction). */
#define BFVM_REDUCTION_ATOMIC 0 /* per-voxel claim + float atomics into the winner's slot (the reference's scheme) */
#define BFVM_REDUCTION_SORTED 1 /* agents sorted by voxel at upload, one segmented sum per voxel, no atomics */

/* Sweep algorithm (bfvm_set_sweep_algorithm). */
#define BFVM_SWEEP_PARTITIONED 0 /* register-resident partitioned (modified) Thomas, FP64-derived coefficients */
#define BFVM_SWEEP_EXACT 1		 /* one thread per line, the reference's FP32 coefficients and operation order */
#define BFVM_SWEEP_STAGED 2		 /* the partitioned algorithm on tiles staged through LDS, one line per thread */

/*
 * Microenvironment description: the fields of `microenvironment` (include/BioFVM/microenvironment.h:19-63)
 * and `cartesian_mesh` (include/BioFVM/mesh.h:7-29) that the solvers read.  Pointers are read during
 * bfvm_initialize only.  A NULL boundary-values pointer means "no condition on that face"
 * (microenvironment.cpp:106-109 default).
 */
typedef struct bfvm_microenvironment_desc
{
	int32_t dims; /* 1, 2 or 3 */
	int32_t bounding_box_mins[3];
	int32_t bounding_box_maxs[3];
	int32_t voxel_shape[3];
	int32_t substrates_count;
	float diffusion_time_step;
	const float* diffusion_coefficients; /* [S] */
	const float* decay_rates;			 /* [S] */

	const float* dirichlet_min_boundary_values[3];		   /* [S] or NULL, per axis */
	const float* dirichlet_max_boundary_values[3];		   /* [S] or NULL */
	const uint8_t* dirichlet_min_boundary_conditions[3]; /* [S] */
	const uint8_t* dirichlet_max_boundary_conditions[3]; /* [S] */

	int32_t dirichlet_interior_voxels_count;
	const int32_t* dirichlet_interior_voxels;		 /* [count][dims] */
	const float* dirichlet_interior_values;		 /* [count][S] */
	const uint8_t* dirichlet_interior_conditions; /* [count][S] */

	int32_t compute_internalized_substrates;
} bfvm_microenvironment_desc;

/* Agent SoA (include/BioFVM/agent_data.h:11-34). fraction_released_at_death may be NULL when unused. */
typedef struct bfvm_agents_desc
{
	int32_t agents_count;
	const float* secretion_rates;			 /* [A][S] */
	const float* saturation_densities;		 /* [A][S] */
	const float* uptake_rates;				 /* [A][S] */
	const float* net_export_rates;			 /* [A][S] */
	const float* internalized_substrates;	 /* [A][S] */
	const float* fraction_released_at_death; /* [A][S] or NULL */
	const float* volumes;					 /* [A] */
	const float* positions;					 /* [A][dims] */
} bfvm_agents_desc;

/* Device addresses of the context's buffers, for harnesses that keep data resident in HBM. */
typedef struct bfvm_device_buffers
{
	void* densities;			   /* [V*S] in the storage form (float for BFVM_STORAGE_F32) */
	float* gradients;			   /* [dims*V*S] */
	float* internalized_substrates; /* [capacity*S] */
	int64_t voxel_count;
	int32_t grid_shape[3];
	int32_t substrates_count;
	void* stream;				 /* hipStream_t the context enqueues on */
	int32_t storage;			 /* BFVM_STORAGE_* of `densities` */
	const float* density_scales; /* [S] device: stored = value * density_scales[s] (all 1 unless FP16) */
} bfvm_device_buffers;

const char* bfvm_last_error(void);

/* device_context ctor: selects `device`, creates the context's stream. */
int bfvm_create(int32_t device, bfvm_ctx** out);
int bfvm_destroy(bfvm_ctx* ctx);

/* solver::initialize: allocates densities (filled with zeros) and gradients, uploads the Dirichlet
 * description, computes the Thomas coefficients on the GPU.  May be called again to re-initialise. */
int bfvm_initialize(bfvm_ctx* ctx, const bfvm_microenvironment_desc* desc);

/* Density storage form (BFVM_STORAGE_*), no counterpart in the reference (its precision switch is the FP64
 * build, CMakeLists.txt:64-66).  Takes effect at the next bfvm_initialize; default FP32.  The 16-bit forms
 * run the partitioned sweep only (the exact sweep reports BFVM_ERR_UNSUPPORTED); densities cross the API in
 * FP32 and are converted on the device. */
int bfvm_set_storage(bfvm_ctx* ctx, int32_t storage);

/* bulk_solver (host/bulk_solver.cpp:23-35) on the device: per voxel and substrate
 * D = (D + dt*S*T) / (1 + dt*(U + S)).  The reference evaluates three per-voxel callbacks at every solve
 * (microenvironment.h:52 supply_rate_func, uptake_rate_func, supply_target_densities_func); here their
 * values are uploaded as [V][S] fields in the densities' layout and kept until the next upload.  The FP32
 * result equals the reference's bit for bit. */
int bfvm_bulk_store_rates(bfvm_ctx* ctx, const float* supply_rates, const float* uptake_rates,
						  const float* supply_target_densities);
int bfvm_bulk_solve(bfvm_ctx* ctx);

/* How agents sharing a voxel are combined (BFVM_REDUCTION_*), no counterpart in the reference, whose cell
 * solvers always claim-and-add (host/cell_solver.cpp:87-135).  The sorted form orders the agents by voxel
 * when they are uploaded (rocPRIM radix sort) and sums each voxel's run in agent order with the lowest agent
 * index as the winner, so it is deterministic and free of contended atomics.  Takes effect at the next
 * bfvm_store_agents; default BFVM_REDUCTION_ATOMIC. */
int bfvm_set_secretion_reduction(bfvm_ctx* ctx, int32_t mode);

/* Selects the sweep algorithm (default BFVM_SWEEP_PARTITIONED) and, for the partitioned one, the
 * maximum chunk length per thread (0 = automatic).  Takes effect at the next solve. */
int bfvm_set_sweep_algorithm(bfvm_ctx* ctx, int32_t algorithm, int32_t chunk);

/* Per-axis tuning of the partitioned sweep: register chunk (8..64, 0 = automatic) and lines per workgroup
 * (y/z: consecutive columns per workgroup, x: rows per workgroup; 0 = automatic).  On a configuration the
 * register/LDS budget cannot hold, the previous one is kept and an error is returned. */
int bfvm_set_sweep_tuning(bfvm_ctx* ctx, int32_t axis, int32_t chunk, int32_t lines_per_group);
/* Most workgroups a partitioned or staged sweep launches (0 = automatic: 8192 partitioned, 4096 staged).
 * Each workgroup loops over line groups (or tiles) in steps of the launched count, so a small cap makes
 * every workgroup run several passes.  No counterpart in the reference, whose sweeps launch one 256-thread
 * block per 256 lines (CUDA/diffusion_solver.cu:947-987). */
int bfvm_set_sweep_workgroups(bfvm_ctx* ctx, int32_t workgroups);
/* Workgroups the 3-D gradient kernel aims for by splitting z into chunks (0 = automatic). */
int bfvm_set_gradient_tuning(bfvm_ctx* ctx, int32_t workgroups);
/* Tile layout of the 3-D gradient: 0 = 256 columns by TY rows per workgroup, 1 = the workgroup's four waves
 * stacked along y (64 columns by 4 TY rows, the rows where waves meet passed through LDS).  Same results. */
int bfvm_set_gradient_stacked(bfvm_ctx* ctx, int32_t stacked);
/* Workgroup order of the 3-D gradient: 0 = each XCD runs consecutive z chunks of one row band, 1 = plane
 * order (the workgroups resident together cover one z chunk of every row band).  Same results.  No
 * counterpart in the reference, whose gradient is a host loop (host/gradient_solver.cpp:275-291). */
int bfvm_set_gradient_order(bfvm_ctx* ctx, int32_t order);

/* Autotuner, in place of the reference's fixed device launch shapes (work_block_size = 30,
 * CUDA/diffusion_solver.cpp:11; 256-thread blocks, CUDA/diffusion_solver.cu:947-987): times candidate shapes
 * of every sweep axis (register chunk, lines or rows per workgroup, workgroup count) and of the 3-D gradient
 * (workgroups, tile layout, workgroup order) on this context's mesh and keeps the fastest, as
 * bfvm_set_sweep_tuning / bfvm_set_sweep_workgroups / bfvm_set_gradient_tuning / bfvm_set_gradient_stacked /
 * bfvm_set_gradient_order would set them.  Results are cached per (axis, mesh
 * shape, substrate count, storage, algorithm) for the process, so a second context of the same shape tunes
 * at no cost.  The candidates run on a device copy of the densities: the densities are unchanged and the
 * gradient buffer is recomputed from them.  Call after bfvm_initialize (and after bfvm_set_storage /
 * bfvm_set_sweep_algorithm have taken effect); synchronous. */
int bfvm_autotune(bfvm_ctx* ctx);
/* The process's tuning cache (bfvm_autotune and bfvm_batch_autotune results) written to / merged from a text
 * file, so that a later process starts tuned: with every shape of a context or batch in the cache, the
 * autotune calls apply the cached shapes without timing anything. */
int bfvm_tuning_save(const char* path);
int bfvm_tuning_load(const char* path);

/* store_data_to_solver / load_data_from_solver parts.  Host pointers, synchronous w.r.t. the host buffer. */
int bfvm_store_densities(bfvm_ctx* ctx, const float* densities);
int bfvm_store_agents(bfvm_ctx* ctx, const bfvm_agents_desc* agents);
int bfvm_load_densities(bfvm_ctx* ctx, float* densities);
int bfvm_load_gradients(bfvm_ctx* ctx, float* gradients);
int bfvm_load_internalized(bfvm_ctx* ctx, float* internalized_substrates);

/* Solvers (asynchronous on the context's stream). */
int bfvm_diffusion_solve(bfvm_ctx* ctx);
int bfvm_dirichlet_solve(bfvm_ctx* ctx);
int bfvm_gradient_solve(bfvm_ctx* ctx);
int bfvm_cell_simulate(bfvm_ctx* ctx, int32_t recompute);
int bfvm_cell_release(bfvm_ctx* ctx, int32_t agent_index);

/* One timestep, secretion -> diffusion (Dirichlet, x, Dirichlet, y, Dirichlet, z, Dirichlet) -> gradient,
 * replayed from a captured HIP graph (captured on first use per (recompute, with_gradient) pair). */
int bfvm_timestep(bfvm_ctx* ctx, int32_t recompute, int32_t with_gradient);

int bfvm_wait_for_all(bfvm_ctx* ctx);

int bfvm_get_device_buffers(bfvm_ctx* ctx, bfvm_device_buffers* out);

/* Device-resident variants for harnesses: same as the host ones but from/to device memory on the
 * context's stream (asynchronous, except that FP16 storage synchronises once to choose its scales). */
int bfvm_store_densities_device(bfvm_ctx* ctx, const float* device_densities);

/*
 * Diagnostics: the per-axis partition the partitioned sweep uses (blocks, block length) and the
 * reference-order FP32 Thomas coefficients b' [n][S], c [S], e [n-1][S] computed on the GPU
 * (host/diffusion_solver.cpp:72-142).
 */
int bfvm_sweep_partition(bfvm_ctx* ctx, int32_t axis, int32_t* blocks, int32_t* block_length,
						 int32_t* lines_per_group);
int bfvm_exact_coefficients(bfvm_ctx* ctx, int32_t axis, float* b, float* c, float* e);

/* mesh::voxel_position<dims> (src/mesh.cpp:47-53) for `count` positions [count][dims], computed by the GPU
 * kernel the secretion solver uses; out3 gets [count][3] indices, -1 on axes where the point is outside. */
int bfvm_voxel_positions(bfvm_ctx* ctx, int32_t count, const float* positions, int32_t* out3);

/* bfvm_voxel_positions without any device allocation, for callers that cannot allocate device memory
 * (code near HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
 * the reference.  The positions and indices are staged through a fixed device array of the library (24 MiB per
 * GPU, held until the process ends) instead of per-call allocations; at most 1,048,576 points per call (more
 * is BFVM_ERR_ARGUMENT).  Every call on one GPU shares that array: calls that overlap on one GPU (from several
 * host threads) race.  Synchronous; results equal bfvm_voxel_positions'. */
int bfvm_voxel_positions_noalloc(bfvm_ctx* ctx, int32_t count, const float* positions, int32_t* out3);

/*
 * Measurement (no counterpart in the reference, whose only timing is std::chrono around whole phases,
 * src/main.cpp:73-101): enqueues `iterations` back-to-back launches of one kernel or phase on the context's
 * stream between two HIP events and returns the average milliseconds per launch.  The work runs on the
 * context's real data (densities are modified).
 */
#define BFVM_KERNEL_SWEEP_X 0
#define BFVM_KERNEL_SWEEP_Y 1
#define BFVM_KERNEL_SWEEP_Z 2
#define BFVM_KERNEL_GRADIENT 3
#define BFVM_KERNEL_DIRICHLET_INTERIOR 4
#define BFVM_KERNEL_DIRICHLET_FACES 5
#define BFVM_KERNEL_COPY 6 /* calibration: float4 copy of the densities into the gradient buffer (8*V*S bytes) */
#define BFVM_PHASE_SECRETION 10		   /* cell solver, recompute = false */
#define BFVM_PHASE_SECRETION_RECOMPUTE 11 /* cell solver, recompute = true */
#define BFVM_PHASE_DIFFUSION 12
#define BFVM_PHASE_GRADIENT 13
#define BFVM_PHASE_TIMESTEP 14 /* graph replay, recompute = false, with gradient */
int bfvm_benchmark(bfvm_ctx* ctx, int32_t what, int32_t iterations, float* avg_ms);

/*
 * Batches of independent microenvironments.  No counterpart in the reference, which runs one
 * microenvironment per solver (src/solver.h); a batch advances many initialized contexts of any mesh shape
 * and substrate count with work-list launches over all instances: per secretion phase, per sweep axis (one
 * launch per thread layout: 4-element vectors where 4 divides the row, else one element) and for the
 * gradient, each driven by a device work list of (instance, block / line-group / tile range) items, and
 * replays the whole step (secretion of every instance -> Dirichlet -> x/y/z sweeps -> gradient) as one
 * captured HIP graph.  Per-instance results equal what bfvm_timestep gives for that context to the
 * partitioned sweep's tolerance (the same algorithm with 16-element blocks).
 *
 * All contexts must be initialized on one device with one storage form and the partitioned sweep; they
 * keep their data (store/load through the per-context calls).  A context re-initialized while the batch
 * exists makes the batch's next step fail with BFVM_ERR_STATE; bfvm_destroy refuses a context that belongs to
 * a batch.  The batch has its own stream, ordered with the contexts' streams both ways: every context call
 * that touches device data first waits for the steps of its batches, and a batch step first waits for work
 * that context calls enqueued on the instances' streams.  bfvm_batch_wait waits for the batch's steps.
 */
typedef struct bfvm_batch bfvm_batch;
int bfvm_batch_create(bfvm_ctx* const* contexts, int32_t count, bfvm_batch** out);
int bfvm_batch_destroy(bfvm_batch* batch);
int bfvm_batch_timestep(bfvm_batch* batch, int32_t recompute, int32_t with_gradient);
int bfvm_batch_wait(bfvm_batch* batch);
/* The batch's autotuner: per sweep axis and thread layout (4-element vectors, one element per thread) the
 * threads a line group aims for and the workgroup count of the work-list launch, and for the gradient the
 * tiles per work item and the workgroup count, each timed on copies of the instances' densities.  The
 * instances' densities are unchanged and their gradients are recomputed.  Synchronous. */
int bfvm_batch_autotune(bfvm_batch* batch);
/* Average ms per launch over all instances: BFVM_KERNEL_SWEEP_X/Y/Z and BFVM_KERNEL_GRADIENT (the batched
 * launch alone), BFVM_PHASE_SECRETION(_RECOMPUTE), _DIFFUSION, _GRADIENT and _TIMESTEP. */
int bfvm_batch_benchmark(bfvm_batch* batch, int32_t what, int32_t iterations, float* avg_ms);

#ifdef __cplusplus
}
#endif

#endif
