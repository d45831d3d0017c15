// C ABI of libparabiofvm_hip.so (declared in include/parabiofvm_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <unordered_map>

#include "bfvm_internal.hpp"

namespace bfvm {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

template <typename F>
static int guarded(const char* fn, F&& f)
{
	try
	{
		return f();
	}
	catch (const hip_failure& e)
	{
		set_error(std::string(fn) + ": HIP error '" + hipGetErrorString(e.err) + "' in " + e.what + " (" + e.file +
				  ":" + std::to_string(e.line) + ")");
		return BFVM_ERR_HIP;
	}
	catch (const std::invalid_argument& e)
	{
		set_error(std::string(fn) + ": " + e.what());
		return BFVM_ERR_ARGUMENT;
	}
	catch (const std::logic_error& e)
	{
		set_error(std::string(fn) + ": " + e.what());
		return BFVM_ERR_STATE;
	}
	catch (const std::exception& e)
	{
		set_error(std::string(fn) + ": " + e.what());
		return BFVM_ERR_UNSUPPORTED;
	}
}

template <typename T>
static void dfree(T*& p)
{
	if (p)
		BFVM_HIP(hipFree(p));
	p = nullptr;
}

template <typename T>
static void dalloc(T*& p, size_t count)
{
	dfree(p);
	if (count > 0)
		BFVM_HIP(hipMalloc(&p, sizeof(T) * count));
}

static void dalloc_bytes(void*& p, size_t bytes)
{
	dfree(p);
	if (bytes > 0)
		BFVM_HIP(hipMalloc(&p, bytes));
}

static void require_init(const bfvm_ctx* ctx)
{
	if (!ctx)
		throw std::invalid_argument("null context");
	if (!ctx->initialized)
		throw std::logic_error("solver not initialized (call bfvm_initialize first)");
}

// Every context entry point that touches device data runs this first: work of a batch holding the context
// (on the batch's own stream) completes before the context reads or rewrites its buffers, and the next batch
// step waits for whatever this call enqueues on the context's stream.
static void enter_ctx(bfvm_ctx* ctx)
{
	for (bfvm_batch* b : ctx->batches)
		BFVM_HIP(hipStreamSynchronize(b->stream));
	ctx->stream_used = true;
}

static void free_agents(bfvm_ctx* ctx)
{
	auto& a = ctx->ag;
	for (float** p : { &a.sec, &a.sat, &a.upt, &a.nex, &a.intern, &a.frac, &a.vol, &a.pos, &a.num, &a.den, &a.fac,
					   &a.rnum, &a.rden, &a.rfac })
		dfree(*p);
	dfree(a.vox);
	dfree(a.win);
	dfree(a.keys);
	dfree(a.idx);
	dfree(a.snum);
	dfree(a.sden);
	dfree(a.sfac);
	dfree(a.multi);
	if (a.sort_tmp)
		BFVM_HIP(hipFree(a.sort_tmp));
	a.sort_tmp = nullptr;
	a.sort_tmp_bytes = 0;
	a.sorted = false;
	a.capacity = 0;
	a.count = 0;
}

static void free_graphs(bfvm_ctx* ctx)
{
	for (auto& g : ctx->graphs)
	{
		if (g.exec)
			BFVM_HIP(hipGraphExecDestroy(g.exec));
		if (g.graph)
			BFVM_HIP(hipGraphDestroy(g.graph));
		g = graph_slot {};
	}
}

static void free_fields(bfvm_ctx* ctx)
{
	dfree(ctx->dens);
	dfree(ctx->dscale);
	dfree(ctx->dinv);
	dfree(ctx->maxbits);
	dfree(ctx->rescale_factor);
	dfree(ctx->rounding_step);
	dfree(ctx->bulk_num);
	dfree(ctx->bulk_den);
	dfree(ctx->staging);
	ctx->staging_elems = 0;
	dfree(ctx->grad);
	dfree(ctx->d_D);
	dfree(ctx->d_decay);
	dfree(ctx->face_val);
	dfree(ctx->face_cond);
	dfree(ctx->int_vox);
	dfree(ctx->int_val);
	dfree(ctx->int_cond);
	dfree(ctx->ballots);
	dfree(ctx->conflict);
	for (auto& t : ctx->axes)
	{
		dfree(t.exact_b);
		dfree(t.exact_c);
		dfree(t.exact_e);
		dfree(t.coef);
		dfree(t.red);
		dfree(t.A);
		t = axis_tables {};
	}
	ctx->n_interior = 0;
	ctx->has_faces = false;
	ctx->initialized = false;
}

static void enqueue_dirichlet(bfvm_ctx* ctx)
{
	launch_dirichlet_faces(ctx);
	launch_dirichlet_interior(ctx);
}

// host/diffusion_solver.cpp:482-544: Dirichlet, x, Dirichlet, y, Dirichlet, z, Dirichlet.  The boundary
// faces of the Dirichlet that follows each sweep are applied in that sweep's store epilogue.
static void enqueue_diffusion(bfvm_ctx* ctx)
{
	if (ctx->algorithm == BFVM_SWEEP_EXACT && ctx->storage != BFVM_STORAGE_F32)
		throw std::runtime_error("the exact sweep needs FP32 density storage");
	if (ctx->storage != BFVM_STORAGE_F32)
		launch_advance_rounding(ctx);
	enqueue_dirichlet(ctx);
	for (int axis = 0; axis < ctx->mesh.dims; axis++)
	{
		launch_sweep(ctx, axis);
		launch_dirichlet_interior(ctx);
	}
}

static void check_gradient(const bfvm_ctx* ctx)
{
	for (int d = 0; d < ctx->mesh.dims; d++)
		if (ctx->mesh.grid[d] < 2)
			throw std::invalid_argument("gradient needs at least 2 voxels along every axis");
}

// Average ms of `iterations` calls of `one` (enqueueing on `stream`), timed with HIP events on that stream
// after one untimed call (graph capture, code object loading).
template <typename F>
static float time_launches(hipStream_t stream, int iterations, F&& one)
{
	one();
	hipEvent_t start = nullptr, stop = nullptr;
	BFVM_HIP(hipEventCreate(&start));
	hipError_t err = hipEventCreate(&stop);
	if (err != hipSuccess)
	{
		(void)hipEventDestroy(start);
		BFVM_HIP(err);
	}
	float ms = 0;
	try
	{
		BFVM_HIP(hipEventRecord(start, stream));
		for (int i = 0; i < iterations; i++)
			one();
		BFVM_HIP(hipEventRecord(stop, stream));
		BFVM_HIP(hipEventSynchronize(stop));
		BFVM_HIP(hipEventElapsedTime(&ms, start, stop));
	}
	catch (...)
	{
		(void)hipEventDestroy(start);
		(void)hipEventDestroy(stop);
		throw;
	}
	BFVM_HIP(hipEventDestroy(start));
	BFVM_HIP(hipEventDestroy(stop));
	return ms / iterations;
}

static void ensure_partitions(bfvm_ctx* ctx)
{
	for (int axis = 0; axis < ctx->mesh.dims; axis++)
	{
		build_partition(ctx, axis, ctx->chunk_request[axis], ctx->lines_request[axis]);
		launch_partition_coefficients(ctx, axis);
	}
}

// FP16 storage: when the bound on |density| of a substrate, times its scale, reaches 2^14 (4x below FP16's
// largest finite value), the stored values and the scale of that substrate shrink by the power of two that
// puts the bound back in [2^10, 2^11).  Enqueued on `stream` ahead of the work that could grow the values; a
// no-op for the other storage forms and while the bound stays below the threshold.
static void ensure_f16_headroom(bfvm_ctx* ctx, hipStream_t stream)
{
	if (ctx->storage != BFVM_STORAGE_F16)
		return;
	const int S = ctx->mesh.S;
	std::vector<float> factor(S, 1.f);
	bool any = false;
	for (int s = 0; s < S; s++)
	{
		const double b = ctx->value_bound[s] * ctx->h_scale[s];
		if (!(b >= 16384.0) || !std::isfinite(ctx->value_bound[s]))
			continue;
		int e;
		std::frexp(ctx->value_bound[s], &e);
		const float scale = std::ldexp(1.f, std::min(100, std::max(-100, 11 - e)));
		if (scale < ctx->h_scale[s])
		{
			factor[s] = scale / ctx->h_scale[s];
			ctx->h_scale[s] = scale;
			any = true;
		}
	}
	if (!any)
		return;
	std::vector<float> inv(S);
	for (int s = 0; s < S; s++)
		inv[s] = 1.f / ctx->h_scale[s];
	BFVM_HIP(hipMemcpyAsync(ctx->rescale_factor, factor.data(), sizeof(float) * S, hipMemcpyHostToDevice, stream));
	launch_rescale_f16(ctx, ctx->rescale_factor, stream);
	BFVM_HIP(hipMemcpyAsync(ctx->dscale, ctx->h_scale.data(), sizeof(float) * S, hipMemcpyHostToDevice, stream));
	BFVM_HIP(hipMemcpyAsync(ctx->dinv, inv.data(), sizeof(float) * S, hipMemcpyHostToDevice, stream));
	BFVM_HIP(hipStreamSynchronize(stream)); // the host vectors above must outlive the copies
}

// `calls` secretion calls are about to be enqueued: each may add the agents' net export to a density
static void account_secretion(bfvm_ctx* ctx, int64_t calls, hipStream_t stream)
{
	if (ctx->storage != BFVM_STORAGE_F16 || ctx->ag.count == 0)
		return;
	for (int s = 0; s < ctx->mesh.S; s++)
		ctx->value_bound[s] += (double)calls * ctx->export_step[s];
	ensure_f16_headroom(ctx, stream);
}

} // namespace bfvm

using namespace bfvm;

extern "C" {

const char* bfvm_last_error(void) { return g_error.c_str(); }

int bfvm_create(int32_t device, bfvm_ctx** out)
{
	return guarded("bfvm_create", [&] {
		if (!out)
			throw std::invalid_argument("null output pointer");
		*out = nullptr;
		int count = 0;
		BFVM_HIP(hipGetDeviceCount(&count));
		if (device < 0 || device >= count)
			throw std::invalid_argument("device index out of range");
		BFVM_HIP(hipSetDevice(device));
		setup_noalloc_scratch(device);
		auto* ctx = new bfvm_ctx();
		ctx->device = device;
		hipError_t err = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
		if (err != hipSuccess)
		{
			delete ctx;
			BFVM_HIP(err);
		}
		*out = ctx;
		return BFVM_OK;
	});
}

int bfvm_destroy(bfvm_ctx* ctx)
{
	if (!ctx)
		return BFVM_OK;
	return guarded("bfvm_destroy", [&] {
		if (!ctx->batches.empty())
			throw std::logic_error("the context belongs to a batch: destroy the batch first");
		BFVM_HIP(hipSetDevice(ctx->device));
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		free_graphs(ctx);
		free_agents(ctx);
		free_fields(ctx);
		BFVM_HIP(hipStreamDestroy(ctx->stream));
		delete ctx;
		return BFVM_OK;
	});
}

int bfvm_initialize(bfvm_ctx* ctx, const bfvm_microenvironment_desc* desc)
{
	return guarded("bfvm_initialize", [&] {
		if (!ctx || !desc)
			throw std::invalid_argument("null argument");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		const int dims = desc->dims;
		const int S = desc->substrates_count;
		if (dims < 1 || dims > 3)
			throw std::invalid_argument("dims must be 1, 2 or 3");
		if (S < 1)
			throw std::invalid_argument("substrates_count must be positive");
		if (!desc->diffusion_coefficients || !desc->decay_rates)
			throw std::invalid_argument("diffusion coefficients and decay rates are required");
		free_graphs(ctx);
		free_agents(ctx);
		free_fields(ctx);
		ctx->generation++;

		mesh_info m;
		m.dims = dims;
		m.S = S;
		m.dt = desc->diffusion_time_step;
		for (int d = 0; d < 3; d++)
		{
			m.mins[d] = desc->bounding_box_mins[d];
			m.maxs[d] = desc->bounding_box_maxs[d];
			m.vs[d] = desc->voxel_shape[d];
			m.grid[d] = 1;
		}
		// mesh.cpp:7-25
		for (int d = 0; d < dims; d++)
		{
			if (m.vs[d] <= 0)
				throw std::invalid_argument("voxel_shape must be positive on every mesh axis");
			m.grid[d] = (m.maxs[d] - m.mins[d] + m.vs[d] - 1) / m.vs[d];
			if (m.grid[d] < 1)
				throw std::invalid_argument("empty mesh axis");
		}
		m.V = (size_t)m.grid[0] * m.grid[1] * m.grid[2];
		ctx->mesh = m;
		ctx->D.assign(desc->diffusion_coefficients, desc->diffusion_coefficients + S);
		ctx->decay.assign(desc->decay_rates, desc->decay_rates + S);
		ctx->compute_internalized = desc->compute_internalized_substrates != 0;

		const size_t VS = m.V * S;
		ctx->storage = ctx->storage_request;
		dalloc_bytes(ctx->dens, VS * storage_bytes(ctx->storage));
		BFVM_HIP(hipMemsetAsync(ctx->dens, 0, VS * storage_bytes(ctx->storage), ctx->stream));
		ctx->h_scale.assign(S, 1.f);
		ctx->value_bound.assign(S, 0.0);
		ctx->export_step.assign(S, 0.0);
		ctx->agent_target.assign(S, 0.f);
		ctx->bulk_target.assign(S, 0.f);
		dalloc(ctx->rescale_factor, S);
		dalloc(ctx->dscale, S);
		dalloc(ctx->dinv, S);
		dalloc(ctx->maxbits, S);
		dalloc(ctx->rounding_step, 1);
		BFVM_HIP(hipMemsetAsync(ctx->rounding_step, 0, sizeof(unsigned), ctx->stream));
		BFVM_HIP(hipMemcpy(ctx->dscale, ctx->h_scale.data(), sizeof(float) * S, hipMemcpyHostToDevice));
		BFVM_HIP(hipMemcpy(ctx->dinv, ctx->h_scale.data(), sizeof(float) * S, hipMemcpyHostToDevice));
		ctx->grad_elems = VS * dims;
		dalloc(ctx->grad, ctx->grad_elems);
		BFVM_HIP(hipMemsetAsync(ctx->grad, 0, sizeof(float) * ctx->grad_elems, ctx->stream));
		dalloc(ctx->d_D, S);
		dalloc(ctx->d_decay, S);
		BFVM_HIP(hipMemcpy(ctx->d_D, ctx->D.data(), sizeof(float) * S, hipMemcpyHostToDevice));
		BFVM_HIP(hipMemcpy(ctx->d_decay, ctx->decay.data(), sizeof(float) * S, hipMemcpyHostToDevice));

		// boundary faces (microenvironment.h:47-50)
		std::vector<float> fval(6 * S, 0.f);
		std::vector<uint8_t> fcond(6 * S, 0);
		bool any = false;
		for (int d = 0; d < dims; d++)
			for (int mx = 0; mx < 2; mx++)
			{
				const float* v = mx ? desc->dirichlet_max_boundary_values[d] : desc->dirichlet_min_boundary_values[d];
				const uint8_t* c =
					mx ? desc->dirichlet_max_boundary_conditions[d] : desc->dirichlet_min_boundary_conditions[d];
				if (!v)
					continue;
				if (!c)
					throw std::invalid_argument("boundary values given without conditions");
				for (int s = 0; s < S; s++)
				{
					fval[(2 * d + mx) * S + s] = v[s];
					fcond[(2 * d + mx) * S + s] = c[s] ? 1 : 0;
					any |= c[s] != 0;
				}
			}
		ctx->has_faces = any;
		ctx->bound_max.assign(S, 0.f);
		for (int f = 0; f < 6; f++)
			for (int s = 0; s < S; s++)
				if (fcond[f * S + s])
					ctx->bound_max[s] = std::max(ctx->bound_max[s], std::fabs(fval[f * S + s]));
		dalloc(ctx->face_val, 6 * S);
		dalloc(ctx->face_cond, 6 * S);
		BFVM_HIP(hipMemcpy(ctx->face_val, fval.data(), sizeof(float) * 6 * S, hipMemcpyHostToDevice));
		BFVM_HIP(hipMemcpy(ctx->face_cond, fcond.data(), 6 * S, hipMemcpyHostToDevice));

		// interior list (microenvironment.h:42-45), one entry per voxel: a later entry overrides an earlier
		// one for every substrate whose condition it sets (the reference applies the list in order)
		const int64_t n_in = desc->dirichlet_interior_voxels_count;
		if (n_in < 0)
			throw std::invalid_argument("negative interior Dirichlet count");
		if (n_in > 0 && (!desc->dirichlet_interior_voxels || !desc->dirichlet_interior_values ||
						 !desc->dirichlet_interior_conditions))
			throw std::invalid_argument("interior Dirichlet arrays missing");
		std::vector<int64_t> vox;
		std::vector<float> val;
		std::vector<uint8_t> cond;
		std::unordered_map<int64_t, size_t> where;
		for (int64_t i = 0; i < n_in; i++)
		{
			int c[3] = { 0, 0, 0 };
			for (int d = 0; d < dims; d++)
			{
				c[d] = desc->dirichlet_interior_voxels[i * dims + d];
				if (c[d] < 0 || c[d] >= m.grid[d])
					throw std::invalid_argument("interior Dirichlet voxel outside the mesh");
			}
			const int64_t lin = ((int64_t)c[2] * m.grid[1] + c[1]) * m.grid[0] + c[0];
			auto it = where.find(lin);
			size_t slot;
			if (it == where.end())
			{
				slot = vox.size();
				where.emplace(lin, slot);
				vox.push_back(lin);
				val.resize(val.size() + S, 0.f);
				cond.resize(cond.size() + S, 0);
			}
			else
				slot = it->second;
			for (int s = 0; s < S; s++)
				if (desc->dirichlet_interior_conditions[i * S + s])
				{
					val[slot * S + s] = desc->dirichlet_interior_values[i * S + s];
					cond[slot * S + s] = 1;
				}
		}
		ctx->n_interior = (int64_t)vox.size();
		for (size_t k = 0; k < val.size(); k++)
			if (cond[k])
				ctx->bound_max[k % S] = std::max(ctx->bound_max[k % S], std::fabs(val[k]));
		if (ctx->n_interior > 0)
		{
			dalloc(ctx->int_vox, vox.size());
			dalloc(ctx->int_val, val.size());
			dalloc(ctx->int_cond, cond.size());
			BFVM_HIP(hipMemcpy(ctx->int_vox, vox.data(), sizeof(int64_t) * vox.size(), hipMemcpyHostToDevice));
			BFVM_HIP(hipMemcpy(ctx->int_val, val.data(), sizeof(float) * val.size(), hipMemcpyHostToDevice));
			BFVM_HIP(hipMemcpy(ctx->int_cond, cond.data(), cond.size(), hipMemcpyHostToDevice));
		}

		// cell solver state (host/cell_solver.cpp:366-373): ballots value-initialised to 0
		dalloc(ctx->ballots, m.V);
		BFVM_HIP(hipMemsetAsync(ctx->ballots, 0, sizeof(int) * m.V, ctx->stream));
		dalloc(ctx->conflict, 1);
		BFVM_HIP(hipMemsetAsync(ctx->conflict, 0, sizeof(int), ctx->stream));

		// diffusion coefficients (host/diffusion_solver.cpp:29-58), computed on the GPU
		for (int axis = 0; axis < dims; axis++)
		{
			auto& t = ctx->axes[axis];
			const int n = m.grid[axis];
			t.n = n;
			dalloc(t.exact_b, (size_t)n * S);
			if (n > 1)
			{
				dalloc(t.exact_c, S);
				dalloc(t.exact_e, (size_t)(n - 1) * S);
			}
			launch_exact_coefficients(ctx, axis);
		}
		ensure_partitions(ctx);
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		ctx->initialized = true;
		ctx->epoch++;
		return BFVM_OK;
	});
}

int bfvm_set_storage(bfvm_ctx* ctx, int32_t storage)
{
	return guarded("bfvm_set_storage", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		if (storage != BFVM_STORAGE_F32 && storage != BFVM_STORAGE_BF16 && storage != BFVM_STORAGE_F16)
			throw std::invalid_argument("unknown storage form");
		ctx->storage_request = storage;
		return BFVM_OK;
	});
}

int bfvm_bulk_store_rates(bfvm_ctx* ctx, const float* supply_rates, const float* uptake_rates,
						  const float* supply_target_densities)
{
	return guarded("bfvm_bulk_store_rates", [&] {
		require_init(ctx);
		if (!supply_rates || !uptake_rates || !supply_target_densities)
			throw std::invalid_argument("null rate field");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		const size_t n = ctx->mesh.V * ctx->mesh.S;
		const float dt = ctx->mesh.dt;
		// host/bulk_solver.cpp:30-33 operation order: (D + dt*S*T) / (1 + dt*(U + S))
		std::vector<float> num(n), den(n);
		for (size_t i = 0; i < n; i++)
		{
			num[i] = dt * supply_rates[i] * supply_target_densities[i];
			den[i] = 1 + dt * (uptake_rates[i] + supply_rates[i]);
		}
		if (!ctx->bulk_num)
		{
			dalloc(ctx->bulk_num, n);
			dalloc(ctx->bulk_den, n);
		}
		BFVM_HIP(hipMemcpy(ctx->bulk_num, num.data(), sizeof(float) * n, hipMemcpyHostToDevice));
		BFVM_HIP(hipMemcpy(ctx->bulk_den, den.data(), sizeof(float) * n, hipMemcpyHostToDevice));
		if (ctx->storage == BFVM_STORAGE_F16)
		{
			// a bulk solve is a weighted average of the density, 0 and the target density
			const int S = ctx->mesh.S;
			std::fill(ctx->bulk_target.begin(), ctx->bulk_target.end(), 0.f);
			for (size_t i = 0; i < n; i++)
				if (supply_rates[i] != 0.f && std::isfinite(supply_target_densities[i]))
					ctx->bulk_target[i % S] = std::max(ctx->bulk_target[i % S], std::fabs(supply_target_densities[i]));
			for (int s = 0; s < S; s++)
				ctx->value_bound[s] = std::max(ctx->value_bound[s], (double)ctx->bulk_target[s]);
			ensure_f16_headroom(ctx, ctx->stream);
		}
		return BFVM_OK;
	});
}

int bfvm_bulk_solve(bfvm_ctx* ctx)
{
	return guarded("bfvm_bulk_solve", [&] {
		require_init(ctx);
		if (!ctx->bulk_num)
			throw std::logic_error("no bulk rates uploaded (call bfvm_bulk_store_rates first)");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		launch_bulk(ctx);
		return BFVM_OK;
	});
}

int bfvm_set_secretion_reduction(bfvm_ctx* ctx, int32_t mode)
{
	return guarded("bfvm_set_secretion_reduction", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		if (mode != BFVM_REDUCTION_ATOMIC && mode != BFVM_REDUCTION_SORTED)
			throw std::invalid_argument("unknown reduction mode");
		ctx->reduction = mode;
		return BFVM_OK;
	});
}

int bfvm_set_sweep_algorithm(bfvm_ctx* ctx, int32_t algorithm, int32_t chunk)
{
	return guarded("bfvm_set_sweep_algorithm", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		if (algorithm != BFVM_SWEEP_PARTITIONED && algorithm != BFVM_SWEEP_EXACT && algorithm != BFVM_SWEEP_STAGED)
			throw std::invalid_argument("unknown sweep algorithm");
		if (algorithm == BFVM_SWEEP_STAGED && chunk != 0 && chunk != batch_chunk)
			throw std::invalid_argument("the staged sweep uses 16-element blocks (chunk 0 or 16)");
		if (algorithm == BFVM_SWEEP_STAGED)
			chunk = batch_chunk;
		if (chunk < 0 || chunk > 64)
			throw std::invalid_argument("chunk must be in [0, 64]");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		ctx->algorithm = algorithm;
		for (int a = 0; a < 3; a++)
			ctx->chunk_request[a] = chunk;
		if (ctx->initialized)
		{
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
			ensure_partitions(ctx);
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
		}
		ctx->epoch++;
		return BFVM_OK;
	});
}

int bfvm_set_sweep_tuning(bfvm_ctx* ctx, int32_t axis, int32_t chunk, int32_t lines_per_group)
{
	return guarded("bfvm_set_sweep_tuning", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		if (axis < 0 || axis > 2)
			throw std::invalid_argument("axis out of range");
		if (chunk < 0 || chunk > 64 || lines_per_group < 0 || lines_per_group > 1024)
			throw std::invalid_argument("chunk must be in [0, 64] and lines_per_group in [0, 1024]");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		const int old_chunk = ctx->chunk_request[axis], old_lines = ctx->lines_request[axis];
		ctx->chunk_request[axis] = chunk;
		ctx->lines_request[axis] = lines_per_group;
		if (ctx->initialized && axis < ctx->mesh.dims)
		{
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
			try
			{
				build_partition(ctx, axis, chunk, lines_per_group);
			}
			catch (...)
			{
				ctx->chunk_request[axis] = old_chunk;
				ctx->lines_request[axis] = old_lines;
				build_partition(ctx, axis, old_chunk, old_lines);
				launch_partition_coefficients(ctx, axis);
				BFVM_HIP(hipStreamSynchronize(ctx->stream));
				throw;
			}
			launch_partition_coefficients(ctx, axis);
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
		}
		ctx->epoch++;
		return BFVM_OK;
	});
}

int bfvm_set_sweep_workgroups(bfvm_ctx* ctx, int32_t workgroups)
{
	return guarded("bfvm_set_sweep_workgroups", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		if (workgroups < 0)
			throw std::invalid_argument("workgroups must be non-negative");
		for (int a = 0; a < 3; a++)
			ctx->sweep_workgroups[a] = workgroups;
		ctx->epoch++;
		return BFVM_OK;
	});
}

int bfvm_set_gradient_tuning(bfvm_ctx* ctx, int32_t workgroups)
{
	return guarded("bfvm_set_gradient_tuning", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		if (workgroups < 0)
			throw std::invalid_argument("workgroups must be non-negative");
		ctx->gradient_groups = workgroups;
		ctx->epoch++;
		return BFVM_OK;
	});
}

int bfvm_set_gradient_stacked(bfvm_ctx* ctx, int32_t stacked)
{
	return guarded("bfvm_set_gradient_stacked", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		ctx->gradient_stacked = stacked != 0;
		ctx->epoch++;
		return BFVM_OK;
	});
}

int bfvm_set_gradient_order(bfvm_ctx* ctx, int32_t order)
{
	return guarded("bfvm_set_gradient_order", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		if (order != 0 && order != 1)
			throw std::invalid_argument("gradient order must be 0 or 1");
		ctx->gradient_order = order;
		ctx->epoch++;
		return BFVM_OK;
	});
}

int bfvm_autotune(bfvm_ctx* ctx)
{
	return guarded("bfvm_autotune", [&] {
		require_init(ctx);
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		bool gradient = ctx->mesh.dims == 3;
		for (int d = 0; d < ctx->mesh.dims; d++)
			gradient &= ctx->mesh.grid[d] >= 2;
		// the candidates run on a copy of the densities, so the context's data stay as they are
		const size_t bytes = ctx->mesh.V * ctx->mesh.S * storage_bytes(ctx->storage);
		void* scratch = nullptr;
		BFVM_HIP(hipMalloc(&scratch, bytes));
		void* real = ctx->dens;
		try
		{
			BFVM_HIP(hipMemcpyAsync(scratch, real, bytes, hipMemcpyDeviceToDevice, ctx->stream));
			ctx->dens = scratch;
			autotune_context(ctx, gradient);
		}
		catch (...)
		{
			ctx->dens = real;
			(void)hipStreamSynchronize(ctx->stream);
			(void)hipFree(scratch);
			throw;
		}
		ctx->dens = real;
		BFVM_HIP(hipFree(scratch));
		// the gradient timings wrote the gradients of the copy: recompute them from the context's densities
		if (gradient)
			launch_gradient(ctx);
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		return BFVM_OK;
	});
}

// FP16 scale of substrate s: the power of two 2^k with (largest |value|) * 2^k in [2^10, 2^11); 1 when the
// substrate is all zeros.  BF16 and FP32 keep 1.
static void choose_scales(bfvm_ctx* ctx, const std::vector<float>& max_abs)
{
	const int S = ctx->mesh.S;
	for (int s = 0; s < S; s++)
	{
		float scale = 1.f;
		const float mx = std::max({ max_abs[s], ctx->bound_max[s], ctx->agent_target[s], ctx->bulk_target[s] });
		ctx->value_bound[s] = mx;
		if (ctx->storage == BFVM_STORAGE_F16 && mx > 0.f && std::isfinite(mx))
		{
			int e;
			std::frexp(mx, &e); // mx = f * 2^e, f in [0.5, 1)
			scale = std::ldexp(1.f, std::min(100, std::max(-100, 11 - e)));
		}
		ctx->h_scale[s] = scale;
	}
	std::vector<float> inv(S);
	for (int s = 0; s < S; s++)
		inv[s] = 1.f / ctx->h_scale[s];
	BFVM_HIP(hipMemcpyAsync(ctx->dscale, ctx->h_scale.data(), sizeof(float) * S, hipMemcpyHostToDevice, ctx->stream));
	BFVM_HIP(hipMemcpyAsync(ctx->dinv, inv.data(), sizeof(float) * S, hipMemcpyHostToDevice, ctx->stream));
	BFVM_HIP(hipStreamSynchronize(ctx->stream));
}

static constexpr size_t staging_chunk = (size_t)1 << 25; // FP32 elements per conversion pass (128 MiB)

static void ensure_staging(bfvm_ctx* ctx)
{
	const size_t want = std::min(staging_chunk, ctx->mesh.V * (size_t)ctx->mesh.S);
	if (ctx->staging_elems < want)
	{
		dalloc(ctx->staging, want);
		ctx->staging_elems = want;
	}
}

int bfvm_store_densities(bfvm_ctx* ctx, const float* densities)
{
	return guarded("bfvm_store_densities", [&] {
		require_init(ctx);
		if (!densities)
			throw std::invalid_argument("null densities");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		const size_t n = ctx->mesh.V * ctx->mesh.S;
		if (ctx->storage == BFVM_STORAGE_F32)
		{
			BFVM_HIP(hipMemcpyAsync(ctx->dens, densities, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
			return BFVM_OK;
		}
		const int S = ctx->mesh.S;
		std::vector<float> mx(S, 0.f);
		if (ctx->storage == BFVM_STORAGE_F16)
			for (size_t i = 0; i < n; i++)
			{
				const float a = std::fabs(densities[i]);
				if (a > mx[i % S] && std::isfinite(a))
					mx[i % S] = a;
			}
		choose_scales(ctx, mx);
		ensure_staging(ctx);
		for (size_t off = 0; off < n; off += ctx->staging_elems)
		{
			const size_t len = std::min(ctx->staging_elems, n - off);
			BFVM_HIP(hipMemcpyAsync(ctx->staging, densities + off, sizeof(float) * len, hipMemcpyHostToDevice,
									ctx->stream));
			launch_pack(ctx, ctx->staging, (int64_t)off, (int64_t)len);
		}
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		return BFVM_OK;
	});
}

int bfvm_store_densities_device(bfvm_ctx* ctx, const float* device_densities)
{
	return guarded("bfvm_store_densities_device", [&] {
		require_init(ctx);
		if (!device_densities)
			throw std::invalid_argument("null densities");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		const size_t n = ctx->mesh.V * ctx->mesh.S;
		if (ctx->storage == BFVM_STORAGE_F32)
		{
			BFVM_HIP(hipMemcpyAsync(ctx->dens, device_densities, sizeof(float) * n, hipMemcpyDeviceToDevice,
									ctx->stream));
			return BFVM_OK;
		}
		const int S = ctx->mesh.S;
		std::vector<float> mx(S, 0.f);
		if (ctx->storage == BFVM_STORAGE_F16)
		{
			BFVM_HIP(hipMemsetAsync(ctx->maxbits, 0, sizeof(unsigned) * S, ctx->stream));
			launch_max_abs(ctx, device_densities, (int64_t)n, ctx->maxbits);
			std::vector<unsigned> bits(S);
			BFVM_HIP(hipMemcpyAsync(bits.data(), ctx->maxbits, sizeof(unsigned) * S, hipMemcpyDeviceToHost,
									ctx->stream));
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
			for (int s = 0; s < S; s++)
				std::memcpy(&mx[s], &bits[s], sizeof(float));
		}
		choose_scales(ctx, mx);
		launch_pack(ctx, device_densities, 0, (int64_t)n);
		return BFVM_OK;
	});
}

int bfvm_store_agents(bfvm_ctx* ctx, const bfvm_agents_desc* agents)
{
	return guarded("bfvm_store_agents", [&] {
		require_init(ctx);
		if (!agents || agents->agents_count < 0)
			throw std::invalid_argument("bad agents description");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		const size_t A = (size_t)agents->agents_count;
		const size_t S = ctx->mesh.S;
		const size_t dims = ctx->mesh.dims;
		if (A > 0 && (!agents->secretion_rates || !agents->saturation_densities || !agents->uptake_rates ||
					  !agents->net_export_rates || !agents->internalized_substrates || !agents->volumes ||
					  !agents->positions))
			throw std::invalid_argument("agent arrays missing");
		auto& a = ctx->ag;
		// device_context.cpp:28-54 / cell_solver.cpp:296-327: reallocate on capacity change
		if (A > a.capacity)
		{
			free_agents(ctx);
			a.capacity = A;
			for (float** p : { &a.sec, &a.sat, &a.upt, &a.nex, &a.intern, &a.frac, &a.num, &a.den, &a.fac, &a.rnum,
							   &a.rden, &a.rfac })
			{
				dalloc(*p, A * S);
				BFVM_HIP(hipMemsetAsync(*p, 0, sizeof(float) * A * S, ctx->stream));
			}
			dalloc(a.vol, A);
			dalloc(a.pos, A * dims);
			dalloc(a.vox, A);
			dalloc(a.win, A);
		}
		const bool sorted = ctx->reduction == BFVM_REDUCTION_SORTED && A > 0;
		if (sorted && !a.keys)
		{
			dalloc(a.keys, 2 * a.capacity);
			dalloc(a.idx, 2 * a.capacity);
			dalloc(a.snum, a.capacity * S);
			dalloc(a.sden, a.capacity * S);
			dalloc(a.sfac, a.capacity * S);
			dalloc(a.multi, 1);
		}
		if (sorted != a.sorted)
			ctx->epoch++;
		a.sorted = sorted;
		if ((int32_t)A != a.count)
			ctx->epoch++;
		a.count = (int32_t)A;
		if (A > 0)
		{
			auto up = [&](float* dst, const float* src, size_t n) {
				BFVM_HIP(hipMemcpyAsync(dst, src, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
			};
			up(a.sec, agents->secretion_rates, A * S);
			up(a.sat, agents->saturation_densities, A * S);
			up(a.upt, agents->uptake_rates, A * S);
			up(a.nex, agents->net_export_rates, A * S);
			up(a.intern, agents->internalized_substrates, A * S);
			if (agents->fraction_released_at_death)
				up(a.frac, agents->fraction_released_at_death, A * S);
			else
				BFVM_HIP(hipMemsetAsync(a.frac, 0, sizeof(float) * A * S, ctx->stream));
			up(a.vol, agents->volumes, A);
			up(a.pos, agents->positions, A * dims);
			// positions change only through this upload: the agent -> voxel table (and the voxel order) is
			// built here
			launch_cell_voxels(ctx);
			if (a.sorted)
				launch_cell_sort(ctx);
		}
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		if (ctx->storage == BFVM_STORAGE_F16)
		{
			// secretion moves a density towards the agents' saturation densities; net export adds N dt / vv per
			// agent and call to the agent's voxel, so a call raises the largest density by at most the largest
			// per-voxel sum of |N| dt / vv (diffusion with decay then never raises a maximum)
			const double vv = (double)ctx->mesh.vs[0] * ctx->mesh.vs[1] * ctx->mesh.vs[2];
			std::fill(ctx->agent_target.begin(), ctx->agent_target.end(), 0.f);
			std::fill(ctx->export_step.begin(), ctx->export_step.end(), 0.0);
			std::unordered_map<int64_t, size_t> slot_of; // voxel -> row of `sums`
			std::vector<double> sums;
			for (size_t i = 0; i < A; i++)
			{
				bool exporting = false;
				for (size_t s = 0; s < S; s++)
				{
					const float t = std::fabs(agents->saturation_densities[i * S + s]);
					if (std::isfinite(t))
						ctx->agent_target[s] = std::max(ctx->agent_target[s], t);
					exporting |= agents->net_export_rates[i * S + s] != 0.f;
				}
				if (!exporting)
					continue;
				// the voxel index of src/mesh.cpp:47-53; agents outside the mesh are ignored by the solver
				int64_t v = 0, mul = 1;
				bool inside = true;
				for (size_t d = 0; d < dims; d++)
				{
					const float q = (agents->positions[i * dims + d] - (float)ctx->mesh.mins[d]) / (float)ctx->mesh.vs[d];
					if (!(q >= 0.f) || !(q < (float)ctx->mesh.grid[d]))
						inside = false;
					else
						v += mul * (int64_t)q;
					mul *= ctx->mesh.grid[d];
				}
				if (!inside)
					continue;
				auto it = slot_of.emplace(v, sums.size() / S).first;
				if (it->second * S == sums.size())
					sums.resize(sums.size() + S, 0.0);
				for (size_t s = 0; s < S; s++)
					sums[it->second * S + s] += std::fabs((double)agents->net_export_rates[i * S + s]);
			}
			for (size_t k = 0; k < sums.size(); k++)
				ctx->export_step[k % S] = std::max(ctx->export_step[k % S], sums[k] * ctx->mesh.dt / vv);
			for (size_t s = 0; s < S; s++)
				ctx->value_bound[s] = std::max(ctx->value_bound[s], (double)ctx->agent_target[s]);
			ensure_f16_headroom(ctx, ctx->stream);
		}
		return BFVM_OK;
	});
}

int bfvm_load_densities(bfvm_ctx* ctx, float* densities)
{
	return guarded("bfvm_load_densities", [&] {
		require_init(ctx);
		if (!densities)
			throw std::invalid_argument("null densities");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		const size_t n = ctx->mesh.V * ctx->mesh.S;
		if (ctx->storage == BFVM_STORAGE_F32)
		{
			BFVM_HIP(hipMemcpyAsync(densities, ctx->dens, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
			return BFVM_OK;
		}
		ensure_staging(ctx);
		for (size_t off = 0; off < n; off += ctx->staging_elems)
		{
			const size_t len = std::min(ctx->staging_elems, n - off);
			launch_unpack(ctx, ctx->staging, (int64_t)off, (int64_t)len);
			BFVM_HIP(hipMemcpyAsync(densities + off, ctx->staging, sizeof(float) * len, hipMemcpyDeviceToHost,
									ctx->stream));
		}
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		return BFVM_OK;
	});
}

int bfvm_load_gradients(bfvm_ctx* ctx, float* gradients)
{
	return guarded("bfvm_load_gradients", [&] {
		require_init(ctx);
		if (!gradients)
			throw std::invalid_argument("null gradients");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		BFVM_HIP(hipMemcpyAsync(gradients, ctx->grad, sizeof(float) * ctx->grad_elems, hipMemcpyDeviceToHost,
								ctx->stream));
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		return BFVM_OK;
	});
}

int bfvm_load_internalized(bfvm_ctx* ctx, float* internalized_substrates)
{
	return guarded("bfvm_load_internalized", [&] {
		require_init(ctx);
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		const size_t n = (size_t)ctx->ag.count * ctx->mesh.S;
		if (n > 0)
		{
			if (!internalized_substrates)
				throw std::invalid_argument("null output");
			BFVM_HIP(hipMemcpyAsync(internalized_substrates, ctx->ag.intern, sizeof(float) * n,
									hipMemcpyDeviceToHost, ctx->stream));
		}
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		return BFVM_OK;
	});
}

int bfvm_diffusion_solve(bfvm_ctx* ctx)
{
	return guarded("bfvm_diffusion_solve", [&] {
		require_init(ctx);
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		enqueue_diffusion(ctx);
		return BFVM_OK;
	});
}

int bfvm_dirichlet_solve(bfvm_ctx* ctx)
{
	return guarded("bfvm_dirichlet_solve", [&] {
		require_init(ctx);
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		enqueue_dirichlet(ctx);
		return BFVM_OK;
	});
}

int bfvm_gradient_solve(bfvm_ctx* ctx)
{
	return guarded("bfvm_gradient_solve", [&] {
		require_init(ctx);
		check_gradient(ctx);
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		launch_gradient(ctx);
		return BFVM_OK;
	});
}

int bfvm_cell_simulate(bfvm_ctx* ctx, int32_t recompute)
{
	return guarded("bfvm_cell_simulate", [&] {
		require_init(ctx);
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		account_secretion(ctx, 1, ctx->stream);
		launch_cell(ctx, recompute != 0);
		return BFVM_OK;
	});
}

int bfvm_cell_release(bfvm_ctx* ctx, int32_t agent_index)
{
	return guarded("bfvm_cell_release", [&] {
		require_init(ctx);
		if (!ctx->compute_internalized)
			return BFVM_OK; // host/cell_solver.cpp:342-343
		if (agent_index < 0 || agent_index >= ctx->ag.count)
			throw std::invalid_argument("agent index out of range");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		if (ctx->storage == BFVM_STORAGE_F16)
		{
			// the release adds I f / vv at the agent's voxel
			const int S = ctx->mesh.S;
			std::vector<float> I(S), f(S);
			BFVM_HIP(hipMemcpyAsync(I.data(), ctx->ag.intern + (size_t)agent_index * S, sizeof(float) * S,
									hipMemcpyDeviceToHost, ctx->stream));
			BFVM_HIP(hipMemcpyAsync(f.data(), ctx->ag.frac + (size_t)agent_index * S, sizeof(float) * S,
									hipMemcpyDeviceToHost, ctx->stream));
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
			const double vv = (double)ctx->mesh.vs[0] * ctx->mesh.vs[1] * ctx->mesh.vs[2];
			for (int s = 0; s < S; s++)
				ctx->value_bound[s] += std::fabs((double)I[s] * f[s]) / vv;
			ensure_f16_headroom(ctx, ctx->stream);
		}
		launch_cell_release(ctx, agent_index);
		return BFVM_OK;
	});
}

int bfvm_timestep(bfvm_ctx* ctx, int32_t recompute, int32_t with_gradient)
{
	return guarded("bfvm_timestep", [&] {
		require_init(ctx);
		if (with_gradient)
			check_gradient(ctx);
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		account_secretion(ctx, 1, ctx->stream);
		auto& slot = ctx->graphs[(recompute ? 2 : 0) + (with_gradient ? 1 : 0)];
		if (!slot.exec || slot.epoch != ctx->epoch)
		{
			if (slot.exec)
				BFVM_HIP(hipGraphExecDestroy(slot.exec));
			if (slot.graph)
				BFVM_HIP(hipGraphDestroy(slot.graph));
			slot = graph_slot {};
			BFVM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
			try
			{
				launch_cell(ctx, recompute != 0);
				enqueue_diffusion(ctx);
				if (with_gradient)
					launch_gradient(ctx);
			}
			catch (...)
			{
				hipGraph_t g = nullptr;
				(void)hipStreamEndCapture(ctx->stream, &g);
				if (g)
					(void)hipGraphDestroy(g);
				throw;
			}
			BFVM_HIP(hipStreamEndCapture(ctx->stream, &slot.graph));
			BFVM_HIP(hipGraphInstantiate(&slot.exec, slot.graph, nullptr, nullptr, 0));
			slot.epoch = ctx->epoch;
		}
		BFVM_HIP(hipGraphLaunch(slot.exec, ctx->stream));
		return BFVM_OK;
	});
}

int bfvm_wait_for_all(bfvm_ctx* ctx)
{
	return guarded("bfvm_wait_for_all", [&] {
		if (!ctx)
			throw std::invalid_argument("null context");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		return BFVM_OK;
	});
}

int bfvm_get_device_buffers(bfvm_ctx* ctx, bfvm_device_buffers* out)
{
	return guarded("bfvm_get_device_buffers", [&] {
		require_init(ctx);
		if (!out)
			throw std::invalid_argument("null output");
		out->densities = ctx->dens;
		out->gradients = ctx->grad;
		out->internalized_substrates = ctx->ag.intern;
		out->voxel_count = (int64_t)ctx->mesh.V;
		for (int d = 0; d < 3; d++)
			out->grid_shape[d] = ctx->mesh.grid[d];
		out->substrates_count = ctx->mesh.S;
		out->stream = (void*)ctx->stream;
		out->storage = ctx->storage;
		out->density_scales = ctx->dscale;
		return BFVM_OK;
	});
}

int bfvm_sweep_partition(bfvm_ctx* ctx, int32_t axis, int32_t* blocks, int32_t* block_length,
						 int32_t* lines_per_group)
{
	return guarded("bfvm_sweep_partition", [&] {
		require_init(ctx);
		if (axis < 0 || axis >= ctx->mesh.dims)
			throw std::invalid_argument("axis out of range");
		const auto& p = ctx->axes[axis].part;
		if (blocks)
			*blocks = p.P;
		if (block_length)
			*block_length = p.P ? p.base + (p.rem > 0 ? 1 : 0) : 1;
		if (lines_per_group)
			*lines_per_group = p.lpg;
		return BFVM_OK;
	});
}

int bfvm_exact_coefficients(bfvm_ctx* ctx, int32_t axis, float* b, float* c, float* e)
{
	return guarded("bfvm_exact_coefficients", [&] {
		require_init(ctx);
		if (axis < 0 || axis >= ctx->mesh.dims)
			throw std::invalid_argument("axis out of range");
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		BFVM_HIP(hipStreamSynchronize(ctx->stream));
		const auto& t = ctx->axes[axis];
		const size_t S = ctx->mesh.S;
		if (b)
			BFVM_HIP(hipMemcpy(b, t.exact_b, sizeof(float) * t.n * S, hipMemcpyDeviceToHost));
		if (c && t.n > 1)
			BFVM_HIP(hipMemcpy(c, t.exact_c, sizeof(float) * S, hipMemcpyDeviceToHost));
		if (e && t.n > 1)
			BFVM_HIP(hipMemcpy(e, t.exact_e, sizeof(float) * (t.n - 1) * S, hipMemcpyDeviceToHost));
		return BFVM_OK;
	});
}

int bfvm_voxel_positions(bfvm_ctx* ctx, int32_t count, const float* positions, int32_t* out3)
{
	return guarded("bfvm_voxel_positions", [&] {
		require_init(ctx);
		if (count < 0 || (count > 0 && (!positions || !out3)))
			throw std::invalid_argument("bad arguments");
		if (count == 0)
			return BFVM_OK;
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		float* d_pos = nullptr;
		int* d_out = nullptr;
		BFVM_HIP(hipMalloc(&d_pos, sizeof(float) * count * ctx->mesh.dims));
		hipError_t err = hipMalloc(&d_out, sizeof(int) * count * 3);
		if (err != hipSuccess)
		{
			(void)hipFree(d_pos);
			BFVM_HIP(err);
		}
		try
		{
			BFVM_HIP(hipMemcpy(d_pos, positions, sizeof(float) * count * ctx->mesh.dims, hipMemcpyHostToDevice));
			launch_voxel_positions(ctx, count, d_pos, d_out);
			BFVM_HIP(hipStreamSynchronize(ctx->stream));
			BFVM_HIP(hipMemcpy(out3, d_out, sizeof(int) * count * 3, hipMemcpyDeviceToHost));
		}
		catch (...)
		{
			(void)hipFree(d_pos);
			(void)hipFree(d_out);
			throw;
		}
		BFVM_HIP(hipFree(d_pos));
		BFVM_HIP(hipFree(d_out));
		return BFVM_OK;
	});
}

int bfvm_voxel_positions_noalloc(bfvm_ctx* ctx, int32_t count, const float* positions, int32_t* out3)
{
	return guarded("bfvm_voxel_positions_noalloc", [&] {
		require_init(ctx);
		if (count < 0 || (count > 0 && (!positions || !out3)))
			throw std::invalid_argument("bad arguments");
		if (count > noalloc_max_points)
			throw std::invalid_argument("count " + std::to_string(count) + " exceeds the " +
										std::to_string(noalloc_max_points) +
										" points of the no-allocation scratch (use bfvm_voxel_positions)");
		if (count == 0)
			return BFVM_OK;
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		voxel_positions_noalloc(ctx, count, positions, out3);
		return BFVM_OK;
	});
}

int bfvm_benchmark(bfvm_ctx* ctx, int32_t what, int32_t iterations, float* avg_ms)
{
	return guarded("bfvm_benchmark", [&] {
		require_init(ctx);
		if (iterations < 1 || !avg_ms)
			throw std::invalid_argument("iterations must be positive and avg_ms non-null");
		const int dims = ctx->mesh.dims;
		if ((what == BFVM_KERNEL_SWEEP_Y && dims < 2) || (what == BFVM_KERNEL_SWEEP_Z && dims < 3))
			throw std::invalid_argument("the mesh has no such axis");
		if (what == BFVM_KERNEL_GRADIENT || what == BFVM_PHASE_GRADIENT || what == BFVM_PHASE_TIMESTEP)
			check_gradient(ctx);
		BFVM_HIP(hipSetDevice(ctx->device));
		enter_ctx(ctx);
		if (what == BFVM_PHASE_SECRETION || what == BFVM_PHASE_SECRETION_RECOMPUTE)
			account_secretion(ctx, (int64_t)iterations + 1, ctx->stream);
		auto one = [&]() {
			switch (what)
			{
			case BFVM_KERNEL_SWEEP_X:
			case BFVM_KERNEL_SWEEP_Y:
			case BFVM_KERNEL_SWEEP_Z:
				launch_sweep(ctx, what - BFVM_KERNEL_SWEEP_X);
				break;
			case BFVM_KERNEL_GRADIENT:
			case BFVM_PHASE_GRADIENT:
				launch_gradient(ctx);
				break;
			case BFVM_KERNEL_DIRICHLET_INTERIOR:
				launch_dirichlet_interior(ctx);
				break;
			case BFVM_KERNEL_DIRICHLET_FACES:
				launch_dirichlet_faces(ctx);
				break;
			case BFVM_KERNEL_COPY:
				launch_copy(ctx);
				break;
			case BFVM_PHASE_SECRETION:
				launch_cell(ctx, false);
				break;
			case BFVM_PHASE_SECRETION_RECOMPUTE:
				launch_cell(ctx, true);
				break;
			case BFVM_PHASE_DIFFUSION:
				enqueue_diffusion(ctx);
				break;
			case BFVM_PHASE_TIMESTEP:
			{
				int rc = bfvm_timestep(ctx, 0, 1);
				if (rc != BFVM_OK)
					throw std::runtime_error(g_error);
				break;
			}
			default:
				throw std::invalid_argument("unknown kernel or phase id");
			}
		};
		*avg_ms = time_launches(ctx->stream, iterations, one);
		return BFVM_OK;
	});
}

} // extern "C"

// ------------------------------------------------------------------------------------------------
// batches of microenvironments
// ------------------------------------------------------------------------------------------------

namespace bfvm {

// The per-instance kernels of a batch step run on the batch's stream with the batch's rounding counter.
struct batch_override
{
	bfvm_ctx* c;
	hipStream_t stream;
	unsigned* rounding;
	batch_override(bfvm_ctx* ctx, const bfvm_batch* b) : c(ctx), stream(ctx->stream), rounding(ctx->rounding_step)
	{
		c->stream = b->stream;
		c->rounding_step = b->rounding;
	}
	~batch_override()
	{
		c->stream = stream;
		c->rounding_step = rounding;
	}
};

static void free_batch(bfvm_batch* b)
{
	for (auto& g : b->graphs)
	{
		if (g.exec)
			(void)hipGraphExecDestroy(g.exec);
		if (g.graph)
			(void)hipGraphDestroy(g.graph);
		g = graph_slot {};
	}
	for (int a = 0; a < 3; a++)
		free_batch_sweep(b, a);
	free_batch_gradient(b);
	free_batch_cell(b);
	if (b->rounding)
		(void)hipFree(b->rounding);
	b->rounding = nullptr;
	if (b->stream)
		(void)hipStreamDestroy(b->stream);
	b->stream = nullptr;
}

static void check_batch(const bfvm_batch* b)
{
	if (!b)
		throw std::invalid_argument("null batch");
	for (size_t i = 0; i < b->ctx.size(); i++)
		if (!b->ctx[i]->initialized || b->ctx[i]->generation != b->generation[i])
			throw std::logic_error("instance " + std::to_string(i) +
								   " was re-initialized after bfvm_batch_create; create the batch again");
}

// A batch step reads every instance's buffers on the batch's stream: first let the work that context calls
// enqueued on the instances' own streams finish.
static void join_instance_streams(bfvm_batch* b)
{
	for (bfvm_ctx* c : b->ctx)
		if (c->stream_used)
		{
			BFVM_HIP(hipStreamSynchronize(c->stream));
			c->stream_used = false;
		}
}

static uint64_t batch_epoch(const bfvm_batch* b)
{
	uint64_t e = b->epoch;
	for (const bfvm_ctx* c : b->ctx)
		e += c->epoch;
	return e;
}

static void enqueue_batch_sweeps(bfvm_batch* b)
{
	if (b->storage != BFVM_STORAGE_F32)
	{
		batch_override o(b->ctx[0], b);
		launch_advance_rounding(b->ctx[0]);
	}
	for (bfvm_ctx* c : b->ctx)
	{
		batch_override o(c, b);
		enqueue_dirichlet(c);
	}
	for (int axis = 0; axis < 3; axis++)
	{
		for (int i : b->axes[axis].single)
		{
			batch_override o(b->ctx[i], b);
			launch_sweep(b->ctx[i], axis);
		}
		launch_batched_sweep(b, axis);
		for (bfvm_ctx* c : b->ctx)
			if (axis < c->mesh.dims)
			{
				batch_override o(c, b);
				launch_dirichlet_interior(c);
			}
	}
}

static void enqueue_batch_gradient(bfvm_batch* b)
{
	launch_batched_gradient(b);
	for (int i : b->grad.single)
	{
		batch_override o(b->ctx[i], b);
		launch_gradient(b->ctx[i]);
	}
}

// the secretion of every instance: batched launches when the plan allows, else instance by instance
static void enqueue_batch_cell(bfvm_batch* b, bool recompute)
{
	if (b->cell.batched)
	{
		launch_batched_cell(b, recompute);
		return;
	}
	for (bfvm_ctx* c : b->ctx)
	{
		batch_override o(c, b);
		launch_cell(c, recompute);
	}
}

static void enqueue_batch_step(bfvm_batch* b, bool recompute, bool with_gradient)
{
	enqueue_batch_cell(b, recompute);
	enqueue_batch_sweeps(b);
	if (with_gradient)
		enqueue_batch_gradient(b);
}

} // namespace bfvm

extern "C" {

int bfvm_batch_create(bfvm_ctx* const* contexts, int32_t count, bfvm_batch** out)
{
	return guarded("bfvm_batch_create", [&] {
		if (!contexts || count < 1 || !out)
			throw std::invalid_argument("need a positive count of contexts and an output pointer");
		*out = nullptr;
		auto* b = new bfvm_batch;
		try
		{
			for (int32_t i = 0; i < count; i++)
			{
				bfvm_ctx* c = contexts[i];
				if (!c || !c->initialized)
					throw std::invalid_argument("context " + std::to_string(i) + " is null or not initialized");
				if (c->device != contexts[0]->device || c->storage != contexts[0]->storage)
					throw std::invalid_argument("all contexts of a batch need the same device and storage form");
				if (c->algorithm == BFVM_SWEEP_EXACT)
					throw std::invalid_argument("batched sweeps use the partitioned algorithm");
				for (int32_t j = 0; j < i; j++)
					if (contexts[j] == c)
						throw std::invalid_argument("a context appears twice in the batch");
				b->ctx.push_back(c);
				b->generation.push_back(c->generation);
			}
			b->device = contexts[0]->device;
			b->storage = contexts[0]->storage;
			BFVM_HIP(hipSetDevice(b->device));
			for (bfvm_ctx* c : b->ctx)
			{
				BFVM_HIP(hipStreamSynchronize(c->stream));
				c->stream_used = false;
			}
			BFVM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
			BFVM_HIP(hipMalloc(&b->rounding, sizeof(unsigned)));
			BFVM_HIP(hipMemsetAsync(b->rounding, 0, sizeof(unsigned), b->stream));
			for (int a = 0; a < 3; a++)
				prepare_batch_sweep(b, a);
			prepare_batch_gradient(b);
			BFVM_HIP(hipStreamSynchronize(b->stream));
		}
		catch (...)
		{
			free_batch(b);
			delete b;
			throw;
		}
		for (bfvm_ctx* c : b->ctx)
			c->batches.push_back(b);
		*out = b;
		return BFVM_OK;
	});
}

int bfvm_batch_destroy(bfvm_batch* batch)
{
	return guarded("bfvm_batch_destroy", [&] {
		if (!batch)
			return BFVM_OK;
		BFVM_HIP(hipSetDevice(batch->device));
		if (batch->stream)
			BFVM_HIP(hipStreamSynchronize(batch->stream));
		for (bfvm_ctx* c : batch->ctx)
			c->batches.erase(std::remove(c->batches.begin(), c->batches.end(), batch), c->batches.end());
		free_batch(batch);
		delete batch;
		return BFVM_OK;
	});
}

int bfvm_batch_timestep(bfvm_batch* batch, int32_t recompute, int32_t with_gradient)
{
	return guarded("bfvm_batch_timestep", [&] {
		check_batch(batch);
		if (with_gradient)
			for (const bfvm_ctx* c : batch->ctx)
				check_gradient(c);
		BFVM_HIP(hipSetDevice(batch->device));
		join_instance_streams(batch);
		for (bfvm_ctx* c : batch->ctx)
			account_secretion(c, 1, batch->stream);
		auto& slot = batch->graphs[(recompute ? 2 : 0) + (with_gradient ? 1 : 0)];
		const uint64_t epoch = batch_epoch(batch);
		refresh_batch_cell(batch, epoch);
		if (!slot.exec || slot.epoch != epoch)
		{
			if (slot.exec)
				BFVM_HIP(hipGraphExecDestroy(slot.exec));
			if (slot.graph)
				BFVM_HIP(hipGraphDestroy(slot.graph));
			slot = graph_slot {};
			BFVM_HIP(hipStreamBeginCapture(batch->stream, hipStreamCaptureModeThreadLocal));
			try
			{
				enqueue_batch_step(batch, recompute != 0, with_gradient != 0);
			}
			catch (...)
			{
				hipGraph_t g = nullptr;
				(void)hipStreamEndCapture(batch->stream, &g);
				if (g)
					(void)hipGraphDestroy(g);
				throw;
			}
			BFVM_HIP(hipStreamEndCapture(batch->stream, &slot.graph));
			BFVM_HIP(hipGraphInstantiate(&slot.exec, slot.graph, nullptr, nullptr, 0));
			slot.epoch = epoch;
		}
		BFVM_HIP(hipGraphLaunch(slot.exec, batch->stream));
		return BFVM_OK;
	});
}

int bfvm_batch_autotune(bfvm_batch* batch)
{
	return guarded("bfvm_batch_autotune", [&] {
		check_batch(batch);
		BFVM_HIP(hipSetDevice(batch->device));
		join_instance_streams(batch);
		BFVM_HIP(hipStreamSynchronize(batch->stream));
		if (apply_cached_batch_tuning(batch))
		{
			for (int a = 0; a < 3; a++)
			{
				free_batch_sweep(batch, a);
				prepare_batch_sweep(batch, a);
			}
			free_batch_gradient(batch);
			prepare_batch_gradient(batch);
			BFVM_HIP(hipStreamSynchronize(batch->stream));
			batch->epoch++;
			return BFVM_OK;
		}
		// the candidates run on copies of the instances' densities
		std::vector<void*> real(batch->ctx.size(), nullptr), scratch(batch->ctx.size(), nullptr);
		auto restore = [&]() {
			for (size_t i = 0; i < batch->ctx.size(); i++)
				if (real[i])
					batch->ctx[i]->dens = real[i];
			(void)hipStreamSynchronize(batch->stream);
			for (void* p : scratch)
				if (p)
					(void)hipFree(p);
		};
		try
		{
			for (size_t i = 0; i < batch->ctx.size(); i++)
			{
				bfvm_ctx* c = batch->ctx[i];
				const size_t bytes = c->mesh.V * c->mesh.S * storage_bytes(c->storage);
				BFVM_HIP(hipMalloc(&scratch[i], bytes));
				BFVM_HIP(hipMemcpyAsync(scratch[i], c->dens, bytes, hipMemcpyDeviceToDevice, batch->stream));
				real[i] = c->dens;
				c->dens = scratch[i];
			}
			autotune_batch(batch);
		}
		catch (...)
		{
			restore();
			for (int a = 0; a < 3; a++)
			{
				free_batch_sweep(batch, a);
				prepare_batch_sweep(batch, a);
			}
			free_batch_gradient(batch);
			prepare_batch_gradient(batch);
			throw;
		}
		restore();
		// the plans hold the instances' buffers: rebuild them on the real densities, recompute the gradients the
		// timings overwrote, and let the secretion plan refresh
		for (int a = 0; a < 3; a++)
		{
			free_batch_sweep(batch, a);
			prepare_batch_sweep(batch, a);
		}
		free_batch_gradient(batch);
		prepare_batch_gradient(batch);
		enqueue_batch_gradient(batch);
		BFVM_HIP(hipStreamSynchronize(batch->stream));
		batch->epoch++;
		batch->cell.epoch = ~0ull;
		return BFVM_OK;
	});
}

int bfvm_tuning_save(const char* path)
{
	return guarded("bfvm_tuning_save", [&] {
		if (!path)
			throw std::invalid_argument("null path");
		save_tuning(path);
		return BFVM_OK;
	});
}

int bfvm_tuning_load(const char* path)
{
	return guarded("bfvm_tuning_load", [&] {
		if (!path)
			throw std::invalid_argument("null path");
		load_tuning(path);
		return BFVM_OK;
	});
}

int bfvm_batch_wait(bfvm_batch* batch)
{
	return guarded("bfvm_batch_wait", [&] {
		if (!batch)
			throw std::invalid_argument("null batch");
		BFVM_HIP(hipSetDevice(batch->device));
		BFVM_HIP(hipStreamSynchronize(batch->stream));
		return BFVM_OK;
	});
}

int bfvm_batch_benchmark(bfvm_batch* batch, int32_t what, int32_t iterations, float* avg_ms)
{
	return guarded("bfvm_batch_benchmark", [&] {
		check_batch(batch);
		if (iterations < 1 || !avg_ms)
			throw std::invalid_argument("iterations must be positive and avg_ms non-null");
		if (what == BFVM_KERNEL_GRADIENT || what == BFVM_PHASE_TIMESTEP)
			for (const bfvm_ctx* c : batch->ctx)
				check_gradient(c);
		BFVM_HIP(hipSetDevice(batch->device));
		join_instance_streams(batch);
		if (what == BFVM_PHASE_SECRETION || what == BFVM_PHASE_SECRETION_RECOMPUTE)
			for (bfvm_ctx* c : batch->ctx)
				account_secretion(c, (int64_t)iterations + 1, batch->stream);
		refresh_batch_cell(batch, batch_epoch(batch));
		auto one = [&]() {
			switch (what)
			{
			case BFVM_KERNEL_SWEEP_X:
			case BFVM_KERNEL_SWEEP_Y:
			case BFVM_KERNEL_SWEEP_Z:
				launch_batched_sweep(batch, what - BFVM_KERNEL_SWEEP_X);
				break;
			case BFVM_KERNEL_GRADIENT:
				launch_batched_gradient(batch);
				break;
			case BFVM_PHASE_SECRETION:
			case BFVM_PHASE_SECRETION_RECOMPUTE:
				enqueue_batch_cell(batch, what == BFVM_PHASE_SECRETION_RECOMPUTE);
				break;
			case BFVM_PHASE_DIFFUSION:
				enqueue_batch_sweeps(batch);
				break;
			case BFVM_PHASE_GRADIENT:
				enqueue_batch_gradient(batch);
				break;
			case BFVM_PHASE_TIMESTEP:
			{
				int rc = bfvm_batch_timestep(batch, 0, 1);
				if (rc != BFVM_OK)
					throw std::runtime_error(g_error);
				break;
			}
			default:
				throw std::invalid_argument("unknown kernel or phase id for a batch");
			}
		};
		*avg_ms = time_launches(batch->stream, iterations, one);
		return BFVM_OK;
	});
}

} // extern "C"
