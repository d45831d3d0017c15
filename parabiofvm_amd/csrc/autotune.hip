// Launch-shape autotuner.  The reference fixes its device launch shapes (work_block_size = 30 in
// CUDA/diffusion_solver.cpp:11, 256-thread blocks in CUDA/diffusion_solver.cu:947-987); here the shapes of
// every sweep axis (register chunk, lines per workgroup, workgroup count) and of the gradient (workgroups it
// aims for) are picked by timing candidates on the context's own mesh, and cached per (axis, mesh shape,
// substrate count, storage) for the process, so that contexts of the same shape tune once.  Batches tune their
// work-list launches (group size and workgroup count per axis and thread layout, gradient tiles per work item
// and workgroup count) the same way, cached per list of instance shapes.  The cache can be written to and read
// from a text file, so that a later process starts tuned.
#include <cstdio>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>

#include "bfvm_internal.hpp"

namespace bfvm {

namespace {

struct sweep_shape
{
	int chunk = 0, lines = 0, workgroups = 0;
};

using shape_key = std::tuple<int, int, int, int, int, int, int>; // axis (or -1: gradient), X, Y, Z, S, storage,
																   // algorithm

std::mutex g_cache_mutex;
std::map<shape_key, sweep_shape> g_sweep_cache;
struct gradient_shape
{
	int groups = 0;
	bool stacked = false;
	int order = 0;
};
std::map<shape_key, gradient_shape> g_gradient_cache;
std::map<std::string, bfvm_batch::tuning_t> g_batch_cache; // key: storage and every instance's shape

shape_key key_of(const bfvm_ctx* ctx, int axis)
{
	const auto& m = ctx->mesh;
	return shape_key { axis, m.grid[0], m.grid[1], m.grid[2], m.S, ctx->storage, ctx->algorithm };
}

// Fastest of `reps` launches after one untimed launch, each timed alone with HIP events on `stream`.
template <typename F>
float best_ms(hipStream_t stream, int reps, F&& launch)
{
	launch();
	hipEvent_t start = nullptr, stop = nullptr;
	BFVM_HIP(hipEventCreate(&start));
	hipError_t err = hipEventCreate(&stop);
	if (err != hipSuccess)
	{
		(void)hipEventDestroy(start);
		BFVM_HIP(err);
	}
	float best = 1e30f;
	try
	{
		for (int r = 0; r < reps; r++)
		{
			BFVM_HIP(hipEventRecord(start, stream));
			launch();
			BFVM_HIP(hipEventRecord(stop, stream));
			BFVM_HIP(hipEventSynchronize(stop));
			float ms = 0;
			BFVM_HIP(hipEventElapsedTime(&ms, start, stop));
			best = std::min(best, ms);
		}
	}
	catch (...)
	{
		(void)hipEventDestroy(start);
		(void)hipEventDestroy(stop);
		throw;
	}
	BFVM_HIP(hipEventDestroy(start));
	BFVM_HIP(hipEventDestroy(stop));
	return best;
}

void apply_sweep_shape(bfvm_ctx* ctx, int axis, const sweep_shape& s)
{
	ctx->chunk_request[axis] = s.chunk;
	ctx->lines_request[axis] = s.lines;
	ctx->sweep_workgroups[axis] = s.workgroups;
	build_partition(ctx, axis, s.chunk, s.lines);
	launch_partition_coefficients(ctx, axis);
}

// Candidates: register chunks (16-element blocks hold 4-substrate vectors; 16-bit storage instantiates 16 and
// 32 only), lines (y, z) or rows (x) per workgroup, then the workgroup count of the best of those.
sweep_shape search_sweep(bfvm_ctx* ctx, int axis, int reps)
{
	std::vector<int> chunks = { 16, 32 };
	if (ctx->storage == BFVM_STORAGE_F32)
		chunks.push_back(64);
	const std::vector<int> lines =
		axis == 0 ? std::vector<int> { 0, 1, 2, 4, 8 } : std::vector<int> { 0, 4, 8, 16, 24, 32 };
	sweep_shape best;
	float best_t = 1e30f;
	for (int c : chunks)
		for (int l : lines)
		{
			const sweep_shape cand { c, l, 0 };
			try
			{
				apply_sweep_shape(ctx, axis, cand);
			}
			catch (const std::runtime_error&)
			{
				continue; // does not fit the register or LDS budget
			}
			const float t = best_ms(ctx->stream, reps, [&] { launch_sweep(ctx, axis); });
			if (t < best_t)
			{
				best_t = t;
				best = cand;
			}
		}
	if (best_t >= 1e30f)
		throw std::runtime_error("autotune: no sweep shape fits axis " + std::to_string(axis));
	for (int w : { 1024, 2048, 4096, 16384, 65536 })
	{
		sweep_shape cand = best;
		cand.workgroups = w;
		apply_sweep_shape(ctx, axis, cand);
		const float t = best_ms(ctx->stream, reps, [&] { launch_sweep(ctx, axis); });
		if (t < best_t * 0.99f)
		{
			best_t = t;
			best = cand;
		}
	}
	return best;
}

std::string batch_key(const bfvm_batch* b)
{
	std::ostringstream k;
	k << b->storage << ':' << b->ctx.size();
	for (const bfvm_ctx* c : b->ctx)
		k << ';' << c->mesh.grid[0] << ',' << c->mesh.grid[1] << ',' << c->mesh.grid[2] << ',' << c->mesh.S << ','
		  << c->mesh.dims;
	return k.str();
}

} // namespace

bool apply_cached_batch_tuning(bfvm_batch* b)
{
	std::lock_guard<std::mutex> lock(g_cache_mutex);
	auto it = g_batch_cache.find(batch_key(b));
	if (it == g_batch_cache.end())
		return false;
	b->tuning = it->second;
	return true;
}

// Text file, one entry per line:
//   sweep <axis> <X> <Y> <Z> <S> <storage> <algorithm> <chunk> <lines> <workgroups>
//   gradient <X> <Y> <Z> <S> <storage> <algorithm> <workgroups> <stacked> [<order>]
//   batch <key> <group threads [3][2]> <max blocks [3][2]> <gradient tiles> <gradient blocks> <gradient rows>
//         <gradient z planes> <gradient split> <gradient stacked>
void save_tuning(const char* path)
{
	std::lock_guard<std::mutex> lock(g_cache_mutex);
	FILE* f = std::fopen(path, "w");
	if (!f)
		throw std::invalid_argument(std::string("cannot write ") + path);
	for (const auto& [k, v] : g_sweep_cache)
		std::fprintf(f, "sweep %d %d %d %d %d %d %d %d %d %d\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
					 std::get<3>(k), std::get<4>(k), std::get<5>(k), std::get<6>(k), v.chunk, v.lines, v.workgroups);
	for (const auto& [k, v] : g_gradient_cache)
		std::fprintf(f, "gradient %d %d %d %d %d %d %d %d %d\n", std::get<1>(k), std::get<2>(k), std::get<3>(k),
					 std::get<4>(k), std::get<5>(k), std::get<6>(k), v.groups, v.stacked ? 1 : 0, v.order);
	for (const auto& [k, t] : g_batch_cache)
	{
		std::fprintf(f, "batch %s", k.c_str());
		for (int a = 0; a < 3; a++)
			for (int w = 0; w < 2; w++)
				std::fprintf(f, " %d", t.group_threads[a][w]);
		for (int a = 0; a < 3; a++)
			for (int w = 0; w < 2; w++)
				std::fprintf(f, " %d", t.max_blocks[a][w]);
		std::fprintf(f, " %d %d %d %d %d %d\n", t.gradient_tiles, t.gradient_blocks, t.gradient_rows, t.gradient_zchunk,
					 t.gradient_split, t.gradient_stacked);
	}
	const bool bad = std::ferror(f) != 0;
	std::fclose(f);
	if (bad)
		throw std::runtime_error(std::string("error writing ") + path);
}

void load_tuning(const char* path)
{
	FILE* f = std::fopen(path, "r");
	if (!f)
		throw std::invalid_argument(std::string("cannot read ") + path);
	std::string text;
	char buf[4096];
	size_t got;
	while ((got = std::fread(buf, 1, sizeof buf, f)) > 0)
		text.append(buf, got);
	std::fclose(f);
	std::istringstream lines(text);
	std::string line;
	std::lock_guard<std::mutex> lock(g_cache_mutex);
	int lineno = 0;
	while (std::getline(lines, line))
	{
		lineno++;
		if (line.empty())
			continue;
		std::istringstream in(line);
		std::string kind;
		in >> kind;
		bool ok = false;
		if (kind == "sweep")
		{
			int axis, X, Y, Z, S, st, alg;
			sweep_shape v;
			ok = static_cast<bool>(in >> axis >> X >> Y >> Z >> S >> st >> alg >> v.chunk >> v.lines >> v.workgroups);
			if (ok)
				g_sweep_cache[shape_key { axis, X, Y, Z, S, st, alg }] = v;
		}
		else if (kind == "gradient")
		{
			int X, Y, Z, S, st, alg, groups, stacked, order = 0;
			ok = static_cast<bool>(in >> X >> Y >> Z >> S >> st >> alg >> groups >> stacked);
			if (ok && !(in >> order)) // files written before the order field hold 8 numbers
				order = 0;
			ok = ok && (order == 0 || order == 1);
			if (ok)
				g_gradient_cache[shape_key { -1, X, Y, Z, S, st, alg }] = gradient_shape { groups, stacked != 0, order };
		}
		else if (kind == "batch")
		{
			std::string key;
			bfvm_batch::tuning_t t;
			ok = static_cast<bool>(in >> key);
			for (int a = 0; a < 3 && ok; a++)
				for (int w = 0; w < 2 && ok; w++)
					ok = static_cast<bool>(in >> t.group_threads[a][w]);
			for (int a = 0; a < 3 && ok; a++)
				for (int w = 0; w < 2 && ok; w++)
					ok = static_cast<bool>(in >> t.max_blocks[a][w]);
			ok = ok && static_cast<bool>(in >> t.gradient_tiles >> t.gradient_blocks >> t.gradient_rows >>
										  t.gradient_zchunk >> t.gradient_split >> t.gradient_stacked) &&
				 (t.gradient_rows == 4 || t.gradient_rows == 8);
			if (ok)
				g_batch_cache[key] = t;
		}
		if (!ok)
			throw std::invalid_argument(std::string(path) + ":" + std::to_string(lineno) + ": not a tuning entry");
	}
}

void autotune_context(bfvm_ctx* ctx, bool tune_gradient)
{
	const int reps = 3;
	for (int axis = 0; axis < ctx->mesh.dims; axis++)
	{
		if (ctx->mesh.grid[axis] == 1 || ctx->algorithm != BFVM_SWEEP_PARTITIONED)
			continue;
		const shape_key key = key_of(ctx, axis);
		sweep_shape shape;
		bool cached = false;
		{
			std::lock_guard<std::mutex> lock(g_cache_mutex);
			auto it = g_sweep_cache.find(key);
			if (it != g_sweep_cache.end())
			{
				shape = it->second;
				cached = true;
			}
		}
		if (!cached)
			shape = search_sweep(ctx, axis, reps);
		apply_sweep_shape(ctx, axis, shape);
		if (!cached)
		{
			std::lock_guard<std::mutex> lock(g_cache_mutex);
			g_sweep_cache[key] = shape;
		}
	}
	if (tune_gradient && ctx->mesh.dims == 3)
	{
		const shape_key key = key_of(ctx, -1);
		gradient_shape best;
		bool cached = false;
		{
			std::lock_guard<std::mutex> lock(g_cache_mutex);
			auto it = g_gradient_cache.find(key);
			if (it != g_gradient_cache.end())
			{
				best = it->second;
				cached = true;
			}
		}
		if (!cached)
		{
			float best_t = 1e30f;
			for (int order : { 0, 1 })
				for (bool stacked : { false, true })
					for (int g : { 2048, 4096, 8192, 16384, 32768, 65536 })
					{
						ctx->gradient_groups = g;
						ctx->gradient_stacked = stacked;
						ctx->gradient_order = order;
						const float t = best_ms(ctx->stream, reps, [&] { launch_gradient(ctx); });
						if (t < best_t)
						{
							best_t = t;
							best = gradient_shape { g, stacked, order };
						}
					}
			std::lock_guard<std::mutex> lock(g_cache_mutex);
			g_gradient_cache[key] = best;
		}
		ctx->gradient_groups = best.groups;
		ctx->gradient_stacked = best.stacked;
		ctx->gradient_order = best.order;
	}
	BFVM_HIP(hipStreamSynchronize(ctx->stream));
	ctx->epoch++;
}

// Batches: per axis and thread layout the target threads of a line group and the workgroup cap of the
// work-list launch; for the gradient the tiles per work item and the workgroup cap.  Each candidate rebuilds
// that part of the plan.
void autotune_batch(bfvm_batch* b)
{
	const int reps = 3;
	for (int axis = 0; axis < 3; axis++)
	{
		if (!b->axes[axis].any)
			continue;
		for (int w = 0; w < 2; w++)
		{
			if (b->axes[axis].parts[w].nitems == 0)
				continue;
			const int max_threads = w == 0 ? 512 : batch_max_threads;
			int best_threads = b->tuning.group_threads[axis][w], best_blocks = b->tuning.max_blocks[axis][w];
			float best_t = 1e30f;
			for (int threads : { 128, 256, 512, 1024 })
			{
				if (threads > max_threads)
					continue;
				for (int blocks : { 1024, 2048, 4096, 8192 })
				{
					b->tuning.group_threads[axis][w] = threads;
					b->tuning.max_blocks[axis][w] = blocks;
					free_batch_sweep(b, axis);
					try
					{
						prepare_batch_sweep(b, axis);
					}
					catch (const std::runtime_error&)
					{
						continue;
					}
					const float t = best_ms(b->stream, reps, [&] { launch_batched_sweep(b, axis); });
					if (t < best_t)
					{
						best_t = t;
						best_threads = threads;
						best_blocks = blocks;
					}
				}
			}
			b->tuning.group_threads[axis][w] = best_threads;
			b->tuning.max_blocks[axis][w] = best_blocks;
			free_batch_sweep(b, axis);
			prepare_batch_sweep(b, axis);
		}
	}
	if (b->grad.nitems[2] > 0)
	{
		// tile shape first (rows per thread, z planes per tile), then the work distribution
		int best_rows = b->tuning.gradient_rows, best_zc = b->tuning.gradient_zchunk;
		int best_split = b->tuning.gradient_split, best_stacked = b->tuning.gradient_stacked;
		float best_shape = 1e30f;
		for (int stacked = 0; stacked < 2; stacked++)
			for (int split = 0; split < (b->storage == BFVM_STORAGE_F32 ? 2 : 1); split++)
				for (int rows : { 4, 8 })
					for (int zc : { 16, 32, 64 })
					{
						b->tuning.gradient_stacked = stacked;
						b->tuning.gradient_split = split;
						b->tuning.gradient_rows = rows;
						b->tuning.gradient_zchunk = zc;
						free_batch_gradient(b);
						prepare_batch_gradient(b);
						const float t = best_ms(b->stream, reps, [&] { launch_batched_gradient(b); });
						if (t < best_shape)
						{
							best_shape = t;
							best_rows = rows;
							best_zc = zc;
							best_split = split;
							best_stacked = stacked;
						}
					}
		b->tuning.gradient_stacked = best_stacked;
		b->tuning.gradient_split = best_split;
		b->tuning.gradient_rows = best_rows;
		b->tuning.gradient_zchunk = best_zc;
		int best_tiles = b->tuning.gradient_tiles, best_blocks = b->tuning.gradient_blocks;
		float best_t = 1e30f;
		for (int tiles : { 1, 2, 4, 8 })
			for (int blocks : { 2048, 4096, 8192, 16384, 32768, 65536 })
			{
				b->tuning.gradient_tiles = tiles;
				b->tuning.gradient_blocks = blocks;
				free_batch_gradient(b);
				prepare_batch_gradient(b);
				const float t = best_ms(b->stream, reps, [&] { launch_batched_gradient(b); });
				if (t < best_t)
				{
					best_t = t;
					best_tiles = tiles;
					best_blocks = blocks;
				}
			}
		b->tuning.gradient_tiles = best_tiles;
		b->tuning.gradient_blocks = best_blocks;
		free_batch_gradient(b);
		prepare_batch_gradient(b);
	}
	BFVM_HIP(hipStreamSynchronize(b->stream));
	b->epoch++;
	std::lock_guard<std::mutex> lock(g_cache_mutex);
	g_batch_cache[batch_key(b)] = b->tuning;
}

} // namespace bfvm
