// Agent secretion and uptake on gfx950 (host/cell_solver.cpp:15-346; formulas host/cell_solver.h:383-409).
//
//   num = S*T*dt*vol/vv,  den = (U+S)*dt*vol/vv,  fac = N*dt/vv          (per agent and substrate)
//   D  <- (D + sum num) / (1 + sum den) + sum fac                        (per voxel)
//   I  -= vv*((-den*D + num)/(1 + den) + fac)                            (per agent, old D)
//
// Same-voxel contributions are combined per recompute by a claim ("ballot") per voxel: the first agent
// to compare-and-swap the voxel's ballot from -1 wins and every agent adds its terms into the winner's
// reduced slot with float atomics (the winner adds den + 1).  A device flag records whether any voxel had
// a second agent; it selects, on the device, between the fused per-agent update (no conflict) and the
// two-phase update (internalized with the old D for every agent, then the winners update D) — the same
// decision the reference makes on the host (host/cell_solver.cpp:208-244).
//
// Each agent's voxel is computed once when the agents are uploaded (positions change only through an
// upload) and kept in agent_vox; the per-substrate kernels run one thread per (agent, substrate) so that
// consecutive lanes touch consecutive floats.
//
// Agents whose position maps outside the grid (or is not finite) get voxel -1 and are ignored; the
// reference would index out of bounds (src/mesh.cpp:47-53 does not clamp).
#include <rocprim/device/device_radix_sort.hpp>

#include "bfvm_internal.hpp"

namespace bfvm {

struct cell_geom
{
	int dims;
	int mins[3];
	int vs[3];
	int grid[3];
	int S;
};

// mesh.cpp:47-53: (int)((p - (float)min) / (float)dx), IEEE division, truncation toward zero.
// Returns the linear voxel ([z][y][x]) or -1 when outside.
__device__ __forceinline__ int64_t agent_voxel(const cell_geom& g, const float* __restrict__ pos, int64_t i)
{
	int idx[3] = { 0, 0, 0 };
	for (int d = 0; d < g.dims; d++)
	{
		const float q = (pos[i * g.dims + d] - (float)g.mins[d]) / (float)g.vs[d];
		if (!(q > -1.0f) || !(q < (float)g.grid[d]))
			return -1;
		idx[d] = (int)q;
	}
	return ((int64_t)idx[2] * g.grid[1] + idx[1]) * g.grid[0] + idx[0];
}

// agent -> voxel table, at upload
__global__ __launch_bounds__(256) void cell_voxels_kernel(cell_geom g, int A, const float* __restrict__ pos,
														  int64_t* __restrict__ vox)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < A)
		vox[i] = agent_voxel(g, pos, i);
}

// recompute, part 1 (host/cell_solver.cpp:35-85): intermediates and reduced-slot reset per (agent, s);
// ballot reset per agent.
__device__ __forceinline__ void cell_prepare_at(int64_t k, int A, int S, float voxel_volume, float dt,
														   const int64_t* __restrict__ vox,
														   const float* __restrict__ vol, const float* __restrict__ sec,
														   const float* __restrict__ sat, const float* __restrict__ upt,
														   const float* __restrict__ nex, float* __restrict__ num,
														   float* __restrict__ den, float* __restrict__ fac,
														   float* __restrict__ rnum, float* __restrict__ rden,
														   float* __restrict__ rfac, int* __restrict__ ballots,
														   int* __restrict__ conflict)
{
	if (k == 0)
		atomicExch(conflict, 0);
	if (k >= (int64_t)A * S)
		return;
	const int64_t i = k / S;
	const int s = (int)(k - i * S);
	if (s == 0)
	{
		const int64_t v = vox[i];
		if (v >= 0)
			ballots[v] = -1;
	}
	const float cv = vol[i];
	num[k] = sec[k] * sat[k] * dt * cv / voxel_volume;
	den[k] = (upt[k] + sec[k]) * dt * cv / voxel_volume;
	fac[k] = nex[k] * dt / voxel_volume;
	rnum[k] = 0;
	rden[k] = 0;
	rfac[k] = 0;
}

__global__ __launch_bounds__(256) void cell_prepare_kernel(int A, int S, float voxel_volume, float dt,
														   const int64_t* __restrict__ vox,
														   const float* __restrict__ vol, const float* __restrict__ sec,
														   const float* __restrict__ sat, const float* __restrict__ upt,
														   const float* __restrict__ nex, float* __restrict__ num,
														   float* __restrict__ den, float* __restrict__ fac,
														   float* __restrict__ rnum, float* __restrict__ rden,
														   float* __restrict__ rfac, int* __restrict__ ballots,
														   int* __restrict__ conflict)
{
	cell_prepare_at((int64_t)blockIdx.x * blockDim.x + threadIdx.x, A, S, voxel_volume, dt, vox, vol, sec, sat, upt,
					nex, num, den, fac, rnum, rden, rfac, ballots, conflict);
}

// *flag = 1 when `cond` holds for an active lane of the wave: one atomic per wave (by the lowest such lane)
// instead of one per lane, as atomics to one address are serialised
__device__ __forceinline__ void flag_any(int* flag, bool cond)
{
	const unsigned long long m = __ballot(cond);
	if (cond && (unsigned)__lane_id() == (unsigned)(__ffsll((long long)m) - 1))
		atomicOr(flag, 1);
}

// recompute, part 2 (host/cell_solver.cpp:98-122): claim the voxel; win[i] = the agent whose slot collects i
__device__ __forceinline__ void cell_ballot_at(int64_t i, int A, const int64_t* __restrict__ vox,
														  int* __restrict__ ballots, int* __restrict__ win,
														  int* __restrict__ conflict)
{
	if (i >= A)
		return;
	const int64_t v = vox[i];
	if (v < 0)
	{
		win[i] = -1;
		return;
	}
	const int prev = atomicCAS(&ballots[v], -1, (int)i);
	win[i] = prev == -1 ? (int)i : prev;
	flag_any(conflict, prev != -1);
}

__global__ __launch_bounds__(256) void cell_ballot_kernel(int A, const int64_t* __restrict__ vox,
														  int* __restrict__ ballots, int* __restrict__ win,
														  int* __restrict__ conflict)
{
	cell_ballot_at((int64_t)blockIdx.x * blockDim.x + threadIdx.x, A, vox, ballots, win, conflict);
}

// recompute, part 3 (host/cell_solver.cpp:108-133): every (agent, s) adds into its winner's slot; the
// winner adds den + 1.
__device__ __forceinline__ void cell_sum_at(int64_t k, int A, int S, const int* __restrict__ win,
													   const float* __restrict__ num, const float* __restrict__ den,
													   const float* __restrict__ fac, float* __restrict__ rnum,
													   float* __restrict__ rden, float* __restrict__ rfac)
{
	if (k >= (int64_t)A * S)
		return;
	const int64_t i = k / S;
	const int s = (int)(k - i * S);
	const int w = win[i];
	if (w < 0)
		return;
	const int64_t t = (int64_t)w * S + s;
	atomicAdd(&rnum[t], num[k]);
	atomicAdd(&rden[t], w == (int)i ? den[k] + 1 : den[k]);
	atomicAdd(&rfac[t], fac[k]);
}

__global__ __launch_bounds__(256) void cell_sum_kernel(int A, int S, const int* __restrict__ win,
													   const float* __restrict__ num, const float* __restrict__ den,
													   const float* __restrict__ fac, float* __restrict__ rnum,
													   float* __restrict__ rden, float* __restrict__ rfac)
{
	cell_sum_at((int64_t)blockIdx.x * blockDim.x + threadIdx.x, A, S, win, num, den, fac, rnum, rden, rfac);
}

// Sorted reduction.  At upload the agents are radix-sorted by voxel (stable, so a voxel's agents keep
// ascending index order; agents outside the mesh get the key V and sort last).  The first agent of each
// voxel's run is its winner.  At recompute the winner's thread sums the run in that order, starting from
// the winner's own terms with den + 1 as in the claim scheme, and writes the winner's slot directly.
__global__ __launch_bounds__(256) void cell_keys_kernel(int A, const int64_t* __restrict__ vox, uint64_t V,
														uint64_t* __restrict__ keys, int* __restrict__ idx)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= A)
		return;
	keys[i] = vox[i] < 0 ? V : (uint64_t)vox[i];
	idx[i] = (int)i;
}

// win[i] = the first agent of i's run (-1 outside the mesh); *multi = 1 when a run holds more than one agent
// (one atomic per workgroup)
__global__ __launch_bounds__(256) void cell_runs_kernel(int A, uint64_t V, const uint64_t* __restrict__ keys,
														const int* __restrict__ idx, int* __restrict__ win,
														int* __restrict__ multi)
{
	const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	bool many = false;
	if (p < A)
	{
		const uint64_t key = keys[p];
		if (key == V)
			win[idx[p]] = -1;
		else if (p == 0 || keys[p - 1] != key)
		{
			const int w = idx[p];
			int64_t q = p;
			for (; q < A && keys[q] == key; q++)
				win[idx[q]] = w;
			many = q - p > 1;
		}
	}
	if (__syncthreads_or(many) && threadIdx.x == 0)
		atomicOr(multi, 1);
}

// recompute with the sorted order, part 1: the terms gathered into voxel order, one thread per (sorted
// position, s) with one dependent load each, so that the run sums read consecutive memory.  (Summing the runs
// straight from agent order made every step of a run head's loop wait for an index load and then a gather:
// 19.8 ms for 1e7 clustered agents x 16 substrates.)
__global__ __launch_bounds__(256) void cell_sorted_gather_kernel(int A, int S, uint64_t V,
																 const uint64_t* __restrict__ keys,
																 const int* __restrict__ idx,
																 const float* __restrict__ num,
																 const float* __restrict__ den,
																 const float* __restrict__ fac,
																 float* __restrict__ snum, float* __restrict__ sden,
																 float* __restrict__ sfac)
{
	const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (int64_t)A * S)
		return;
	const int64_t p = t / S;
	const int s = (int)(t - p * S);
	if (keys[p] == V)
		return;
	const int64_t k = (int64_t)idx[p] * S + s;
	snum[t] = num[k];
	sden[t] = den[k];
	sfac[t] = fac[k];
}

// part 2: one thread per (sorted position, s); run heads sum their run (winner first with den + 1, then the
// run's agents in ascending index order) and write the winner's slot.
__global__ __launch_bounds__(256) void cell_sorted_sum_kernel(int A, int S, uint64_t V,
															  const uint64_t* __restrict__ keys,
															  const int* __restrict__ idx,
															  const float* __restrict__ snum,
															  const float* __restrict__ sden,
															  const float* __restrict__ sfac, float* __restrict__ rnum,
															  float* __restrict__ rden, float* __restrict__ rfac,
															  int* __restrict__ ballots, const int* __restrict__ multi,
															  int* __restrict__ conflict)
{
	const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	// the voxel order is fixed at upload, and so is whether a voxel holds several agents
	if (t == 0 && *multi)
		atomicOr(conflict, 1);
	if (t >= (int64_t)A * S)
		return;
	const int64_t p = t / S;
	const int s = (int)(t - p * S);
	const uint64_t key = keys[p];
	if (key == V || (p > 0 && keys[p - 1] == key))
		return;
	const int w = idx[p];
	const int64_t kw = (int64_t)w * S + s;
	float sn = 0.f, sd = 0.f, sf = 0.f;
	sn = sn + snum[t];
	sd = sd + (sden[t] + 1);
	sf = sf + sfac[t];
	int64_t q = p + 1;
	for (; q < A && keys[q] == key; q++)
	{
		const int64_t k = q * S + s;
		sn = sn + snum[k];
		sd = sd + sden[k];
		sf = sf + sfac[k];
	}
	rnum[kw] = sn;
	rden[kw] = sd;
	rfac[kw] = sf;
	if (s == 0)
		ballots[key] = w;
}

// Apply, part 1 (host/cell_solver.cpp:137-152, 173-195, 208-234).  One thread per (agent, substrate).
// No conflict and internalized: the fused update of I and D.  Conflict and internalized: I with the
// agent's own terms and the current D (D is updated by part 2 after the kernel boundary).
template <int K>
__device__ __forceinline__ void cell_apply_at(int64_t k, int A, int S, int with_internalized, int voxel_volume,
														 const int64_t* __restrict__ vox, void* __restrict__ dens,
														 const float* __restrict__ scale, const float* __restrict__ inv,
														 const unsigned* __restrict__ rounding,
														 float* __restrict__ intern, const float* __restrict__ num,
														 const float* __restrict__ den, const float* __restrict__ fac,
														 const float* __restrict__ rnum,
														 const float* __restrict__ rden,
														 const float* __restrict__ rfac,
														 const int* __restrict__ conflict)
{
	if (!with_internalized)
		return;
	if (k >= (int64_t)A * S)
		return;
	const int64_t i = k / S;
	const int s = (int)(k - i * S);
	const int64_t v = vox[i];
	if (v < 0)
		return;
	const int64_t at = v * S + s;
	const float vv = (float)voxel_volume;
	const float d = K == store_f32 ? dload<K, 1>(dens, at) : dload<K, 1>(dens, at) * inv[s];
	if (*conflict == 0)
	{
		const float rn = rnum[k], rd = rden[k], rf = rfac[k];
		intern[k] = intern[k] - vv * (((1 - rd) * d + rn) / (rd) + rf);
		const float nd = (d + rn) / rd + rf;
		dstore_sr<K, 1>(dens, at, K == store_f32 ? nd : nd * scale[s], K == store_f32 ? 0u : rounding_key(rounding, 5));
	}
	else
		intern[k] = intern[k] - vv * (((-den[k]) * d + num[k]) / (1 + den[k]) + fac[k]);
}

template <int K>
__global__ __launch_bounds__(256) void cell_apply_kernel(int A, int S, int with_internalized, int voxel_volume,
														 const int64_t* __restrict__ vox, void* __restrict__ dens,
														 const float* __restrict__ scale, const float* __restrict__ inv,
														 const unsigned* __restrict__ rounding,
														 float* __restrict__ intern, const float* __restrict__ num,
														 const float* __restrict__ den, const float* __restrict__ fac,
														 const float* __restrict__ rnum,
														 const float* __restrict__ rden,
														 const float* __restrict__ rfac,
														 const int* __restrict__ conflict)
{
	cell_apply_at<K>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, A, S, with_internalized, voxel_volume, vox, dens,
					 scale, inv, rounding, intern, num, den, fac, rnum, rden, rfac, conflict);
}

// Apply, part 2 (host/cell_solver.cpp:154-171, 236-244): the ballot winner of each voxel updates D with
// the reduced sums.  Skipped when part 1 already did the fused update.
template <int K>
__device__ __forceinline__ void cell_densities_at(int64_t k, int A, int S, int with_internalized,
															 const int64_t* __restrict__ vox, void* __restrict__ dens,
															 const float* __restrict__ scale,
															 const float* __restrict__ inv,
															 const unsigned* __restrict__ rounding,
															 const float* __restrict__ rnum,
															 const float* __restrict__ rden,
															 const float* __restrict__ rfac,
															 const int* __restrict__ ballots,
															 const int* __restrict__ conflict)
{
	if (with_internalized && *conflict == 0)
		return;
	if (k >= (int64_t)A * S)
		return;
	const int64_t i = k / S;
	const int s = (int)(k - i * S);
	const int64_t v = vox[i];
	if (v < 0)
		return;
	if (ballots[v] != (int)i)
		return;
	const int64_t at = v * S + s;
	const float d = K == store_f32 ? dload<K, 1>(dens, at) : dload<K, 1>(dens, at) * inv[s];
	const float nd = (d + rnum[k]) / rden[k] + rfac[k];
	dstore_sr<K, 1>(dens, at, K == store_f32 ? nd : nd * scale[s], K == store_f32 ? 0u : rounding_key(rounding, 6));
}

template <int K>
__global__ __launch_bounds__(256) void cell_densities_kernel(int A, int S, int with_internalized,
															 const int64_t* __restrict__ vox, void* __restrict__ dens,
															 const float* __restrict__ scale,
															 const float* __restrict__ inv,
															 const unsigned* __restrict__ rounding,
															 const float* __restrict__ rnum,
															 const float* __restrict__ rden,
															 const float* __restrict__ rfac,
															 const int* __restrict__ ballots,
															 const int* __restrict__ conflict)
{
	cell_densities_at<K>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, A, S, with_internalized, vox, dens, scale, inv,
						 rounding, rnum, rden, rfac, ballots, conflict);
}

// Many microenvironments (bfvm_batch): one launch per secretion phase over every instance, workgroups walking
// a work list of (instance, block range) items; each block runs the phase's per-element function for its 256
// elements, as the single-instance kernels do.
enum cell_phase
{
	phase_prepare,
	phase_ballot,
	phase_sum,
	phase_apply,
	phase_densities
};

template <int K, int PHASE>
__global__ __launch_bounds__(256) void cell_batched_kernel(const cell_args* __restrict__ inst,
														   const batch_item* __restrict__ items, int nitems)
{
	for (int it = blockIdx.x; it < nitems; it += gridDim.x)
	{
		const batch_item w = items[it];
		const cell_args c = inst[w.instance];
		for (int64_t blk = w.begin; blk < w.end; blk++)
		{
			const int64_t k = blk * blockDim.x + threadIdx.x;
			if constexpr (PHASE == phase_prepare)
				cell_prepare_at(k, c.A, c.S, (float)c.voxel_volume, c.dt, c.vox, c.vol, c.sec, c.sat, c.upt, c.nex, c.num,
								c.den, c.fac, c.rnum, c.rden, c.rfac, c.ballots, c.conflict);
			else if constexpr (PHASE == phase_ballot)
				cell_ballot_at(k, c.A, c.vox, c.ballots, c.win, c.conflict);
			else if constexpr (PHASE == phase_sum)
				cell_sum_at(k, c.A, c.S, c.win, c.num, c.den, c.fac, c.rnum, c.rden, c.rfac);
			else if constexpr (PHASE == phase_apply)
				cell_apply_at<K>(k, c.A, c.S, c.with_internalized, c.voxel_volume, c.vox, c.dens, c.scale, c.inv,
								 c.rounding, c.intern, c.num, c.den, c.fac, c.rnum, c.rden, c.rfac, c.conflict);
			else
				cell_densities_at<K>(k, c.A, c.S, c.with_internalized, c.vox, c.dens, c.scale, c.inv, c.rounding,
									 c.rnum, c.rden, c.rfac, c.ballots, c.conflict);
		}
	}
}

// host/cell_solver.cpp:301-328: D += I*F/vv, I = 0 for one agent.
template <int K>
__global__ void cell_release_kernel(int S, int64_t i, float voxel_volume, const int64_t* __restrict__ vox,
									void* __restrict__ dens, const float* __restrict__ scale,
									const float* __restrict__ inv, const unsigned* __restrict__ rounding,
									float* __restrict__ intern,
									const float* __restrict__ frac)
{
	const int s = threadIdx.x + blockIdx.x * blockDim.x;
	if (s >= S)
		return;
	const int64_t v = vox[i];
	if (v < 0)
		return;
	const int64_t k = i * S + s;
	const int64_t at = v * S + s;
	const float d = K == store_f32 ? dload<K, 1>(dens, at) : dload<K, 1>(dens, at) * inv[s];
	const float nd = d + intern[k] * frac[k] / voxel_volume;
	dstore_sr<K, 1>(dens, at, K == store_f32 ? nd : nd * scale[s], K == store_f32 ? 0u : rounding_key(rounding, 7));
	intern[k] = 0;
}

__global__ void voxel_positions_kernel(cell_geom g, int count, const float* __restrict__ pos, int* __restrict__ out)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count)
		return;
	for (int d = 0; d < 3; d++)
	{
		int r = 0;
		if (d < g.dims)
		{
			const float q = (pos[i * g.dims + d] - (float)g.mins[d]) / (float)g.vs[d];
			r = (q > -1.0f && q < (float)g.grid[d]) ? (int)q : -1;
		}
		out[i * 3 + d] = r;
	}
}

static cell_geom make_geom(const bfvm_ctx* ctx)
{
	cell_geom g;
	g.dims = ctx->mesh.dims;
	for (int d = 0; d < 3; d++)
	{
		g.mins[d] = ctx->mesh.mins[d];
		g.vs[d] = ctx->mesh.vs[d];
		g.grid[d] = ctx->mesh.grid[d];
	}
	g.S = ctx->mesh.S;
	return g;
}

static int voxel_volume(const bfvm_ctx* ctx) { return ctx->mesh.vs[0] * ctx->mesh.vs[1] * ctx->mesh.vs[2]; }

void launch_cell_voxels(bfvm_ctx* ctx)
{
	const auto& a = ctx->ag;
	if (a.count == 0)
		return;
	hipLaunchKernelGGL(cell_voxels_kernel, dim3((a.count + 255) / 256), dim3(256), 0, ctx->stream, make_geom(ctx),
					   a.count, a.pos, a.vox);
	BFVM_HIP(hipGetLastError());
}

// Keys and agent indices in the first half of a.keys / a.idx, sorted into the second half; rocPRIM's
// temporary storage grows on demand.  Radix passes cover only the bits a voxel index (or V) needs.
void launch_cell_sort(bfvm_ctx* ctx)
{
	auto& a = ctx->ag;
	const int A = a.count;
	if (A == 0)
		return;
	const uint64_t V = ctx->mesh.V;
	hipLaunchKernelGGL(cell_keys_kernel, dim3((A + 255) / 256), dim3(256), 0, ctx->stream, A, a.vox, V, a.keys, a.idx);
	BFVM_HIP(hipGetLastError());
	int bits = 1;
	while (bits < 64 && (V >> bits) != 0)
		bits++;
	size_t need = 0;
	BFVM_HIP(rocprim::radix_sort_pairs(nullptr, need, a.keys, a.keys + a.capacity, a.idx, a.idx + a.capacity,
									   (size_t)A, 0, bits, ctx->stream));
	if (need > a.sort_tmp_bytes)
	{
		if (a.sort_tmp)
			BFVM_HIP(hipFree(a.sort_tmp));
		a.sort_tmp = nullptr;
		a.sort_tmp_bytes = 0;
		BFVM_HIP(hipMalloc(&a.sort_tmp, need));
		a.sort_tmp_bytes = need;
	}
	size_t have = a.sort_tmp_bytes;
	BFVM_HIP(rocprim::radix_sort_pairs(a.sort_tmp, have, a.keys, a.keys + a.capacity, a.idx, a.idx + a.capacity,
									   (size_t)A, 0, bits, ctx->stream));
	BFVM_HIP(hipMemsetAsync(a.multi, 0, sizeof(int), ctx->stream));
	hipLaunchKernelGGL(cell_runs_kernel, dim3((A + 255) / 256), dim3(256), 0, ctx->stream, A, V, a.keys + a.capacity,
					   a.idx + a.capacity, a.win, a.multi);
	BFVM_HIP(hipGetLastError());
}

void launch_cell(bfvm_ctx* ctx, bool recompute)
{
	const auto& a = ctx->ag;
	const int A = a.count;
	const int S = ctx->mesh.S;
	const unsigned agent_blocks = (unsigned)((A + 255) / 256);
	const int64_t pairs = (int64_t)A * S;
	const unsigned pair_blocks = (unsigned)((pairs + 255) / 256);
	if (recompute)
	{
		// at least one block so the conflict flag is cleared even without agents
		hipLaunchKernelGGL(cell_prepare_kernel, dim3(pair_blocks > 0 ? pair_blocks : 1), dim3(256), 0, ctx->stream, A,
						   S, (float)voxel_volume(ctx), ctx->mesh.dt, a.vox, a.vol, a.sec, a.sat, a.upt, a.nex, a.num,
						   a.den, a.fac, a.rnum, a.rden, a.rfac, ctx->ballots, ctx->conflict);
		BFVM_HIP(hipGetLastError());
		if (A > 0)
		{
			if (a.sorted)
			{
				hipLaunchKernelGGL(cell_sorted_gather_kernel, dim3(pair_blocks), dim3(256), 0, ctx->stream, A, S,
								   (uint64_t)ctx->mesh.V, a.keys + a.capacity, a.idx + a.capacity, a.num, a.den, a.fac,
								   a.snum, a.sden, a.sfac);
				BFVM_HIP(hipGetLastError());
				hipLaunchKernelGGL(cell_sorted_sum_kernel, dim3(pair_blocks), dim3(256), 0, ctx->stream, A, S,
								   (uint64_t)ctx->mesh.V, a.keys + a.capacity, a.idx + a.capacity, a.snum, a.sden,
								   a.sfac, a.rnum, a.rden, a.rfac, ctx->ballots, a.multi, ctx->conflict);
				BFVM_HIP(hipGetLastError());
			}
			else
			{
				hipLaunchKernelGGL(cell_ballot_kernel, dim3(agent_blocks), dim3(256), 0, ctx->stream, A, a.vox,
								   ctx->ballots, a.win, ctx->conflict);
				BFVM_HIP(hipGetLastError());
				hipLaunchKernelGGL(cell_sum_kernel, dim3(pair_blocks), dim3(256), 0, ctx->stream, A, S, a.win,
								   a.num, a.den, a.fac, a.rnum, a.rden, a.rfac);
				BFVM_HIP(hipGetLastError());
			}
		}
	}
	if (A == 0)
		return;
	by_storage(ctx->storage, [&](auto kind) {
		constexpr int K = decltype(kind)::value;
		hipLaunchKernelGGL(cell_apply_kernel<K>, dim3(pair_blocks), dim3(256), 0, ctx->stream, A, S,
						   (int)ctx->compute_internalized, voxel_volume(ctx), a.vox, ctx->dens, ctx->dscale, ctx->dinv,
						   ctx->rounding_step, a.intern, a.num, a.den, a.fac, a.rnum, a.rden, a.rfac, ctx->conflict);
		BFVM_HIP(hipGetLastError());
		hipLaunchKernelGGL(cell_densities_kernel<K>, dim3(pair_blocks), dim3(256), 0, ctx->stream, A, S,
						   (int)ctx->compute_internalized, a.vox, ctx->dens, ctx->dscale, ctx->dinv, ctx->rounding_step,
						   a.rnum, a.rden,
						   a.rfac, ctx->ballots, ctx->conflict);
		BFVM_HIP(hipGetLastError());
	});
}

void free_batch_cell(bfvm_batch* b)
{
	auto& cp = b->cell;
	if (cp.inst)
		BFVM_HIP(hipFree(cp.inst));
	if (cp.pair_items)
		BFVM_HIP(hipFree(cp.pair_items));
	if (cp.agent_items)
		BFVM_HIP(hipFree(cp.agent_items));
	cp = bfvm_batch::cell_plan {};
}

// Per-instance arguments and work items of the batched secretion, rebuilt when `epoch` (the batch's sum of
// instance epochs, which moves whenever an instance's agents change) differs from the plan's.  Items cover
// blocks of 256 (agent, substrate) pairs (at least one block per instance, so the recompute clears every
// conflict flag) and blocks of 256 agents, 8 blocks per item.
void refresh_batch_cell(bfvm_batch* b, uint64_t epoch)
{
	auto& cp = b->cell;
	if (cp.epoch == epoch)
		return;
	// the per-instance arguments and work items are rewritten below with blocking copies, which the batch's
	// non-blocking stream is not ordered against: steps still in flight must finish first
	BFVM_HIP(hipStreamSynchronize(b->stream));
	cp.batched = true;
	for (const bfvm_ctx* c : b->ctx)
		cp.batched &= !c->ag.sorted;
	if (!cp.batched)
	{
		cp.epoch = epoch;
		return;
	}
	constexpr int64_t per_item = 8;
	std::vector<cell_args> args(b->ctx.size());
	std::vector<batch_item> pairs, agents;
	for (size_t i = 0; i < b->ctx.size(); i++)
	{
		const bfvm_ctx* ctx = b->ctx[i];
		const auto& a = ctx->ag;
		cell_args& c = args[i];
		c.A = a.count;
		c.S = ctx->mesh.S;
		c.with_internalized = ctx->compute_internalized ? 1 : 0;
		c.voxel_volume = voxel_volume(ctx);
		c.dt = ctx->mesh.dt;
		c.vox = a.vox;
		c.vol = a.vol;
		c.sec = a.sec;
		c.sat = a.sat;
		c.upt = a.upt;
		c.nex = a.nex;
		c.num = a.num;
		c.den = a.den;
		c.fac = a.fac;
		c.rnum = a.rnum;
		c.rden = a.rden;
		c.rfac = a.rfac;
		c.intern = a.intern;
		c.ballots = ctx->ballots;
		c.win = a.win;
		c.conflict = ctx->conflict;
		c.dens = ctx->dens;
		c.scale = ctx->dscale;
		c.inv = ctx->dinv;
		c.rounding = b->rounding;
		const int64_t pair_blocks = std::max<int64_t>(1, ((int64_t)c.A * c.S + 255) / 256);
		const int64_t agent_blocks = ((int64_t)c.A + 255) / 256;
		for (int64_t b0 = 0; b0 < pair_blocks; b0 += per_item)
			pairs.push_back(batch_item { (int32_t)i, 0, b0, std::min(pair_blocks, b0 + per_item) });
		for (int64_t b0 = 0; b0 < agent_blocks; b0 += per_item)
			agents.push_back(batch_item { (int32_t)i, 0, b0, std::min(agent_blocks, b0 + per_item) });
	}
	if (!cp.inst)
		BFVM_HIP(hipMalloc(&cp.inst, sizeof(cell_args) * args.size()));
	BFVM_HIP(hipMemcpy(cp.inst, args.data(), sizeof(cell_args) * args.size(), hipMemcpyHostToDevice));
	auto upload = [](const std::vector<batch_item>& v, batch_item*& dst, size_t& cap) {
		if (v.size() > cap)
		{
			if (dst)
				BFVM_HIP(hipFree(dst));
			dst = nullptr;
			BFVM_HIP(hipMalloc(&dst, sizeof(batch_item) * v.size()));
			cap = v.size();
		}
		if (!v.empty())
			BFVM_HIP(hipMemcpy(dst, v.data(), sizeof(batch_item) * v.size(), hipMemcpyHostToDevice));
	};
	upload(pairs, cp.pair_items, cp.pair_cap);
	upload(agents, cp.agent_items, cp.agent_cap);
	cp.npair = (int)pairs.size();
	cp.nagent = (int)agents.size();
	cp.epoch = epoch;
}

template <int K, int PHASE>
static void launch_cell_phase(bfvm_batch* b, const batch_item* items, int n)
{
	if (n == 0)
		return;
	hipLaunchKernelGGL((cell_batched_kernel<K, PHASE>), dim3(std::min(n, 4096)), dim3(256), 0, b->stream,
					   (const cell_args*)b->cell.inst, items, n);
	BFVM_HIP(hipGetLastError());
}

// The secretion of every instance of a refreshed, batched plan: recompute (prepare, ballot, sum) and apply
// (apply, densities), one launch per phase, in the order of launch_cell.
void launch_batched_cell(bfvm_batch* b, bool recompute)
{
	const auto& cp = b->cell;
	by_storage(b->storage, [&](auto kind) {
		constexpr int K = decltype(kind)::value;
		if (recompute)
		{
			launch_cell_phase<K, phase_prepare>(b, cp.pair_items, cp.npair);
			launch_cell_phase<K, phase_ballot>(b, cp.agent_items, cp.nagent);
			launch_cell_phase<K, phase_sum>(b, cp.pair_items, cp.npair);
		}
		launch_cell_phase<K, phase_apply>(b, cp.pair_items, cp.npair);
		launch_cell_phase<K, phase_densities>(b, cp.pair_items, cp.npair);
	});
}

void launch_cell_release(bfvm_ctx* ctx, int index)
{
	const auto& a = ctx->ag;
	const int S = ctx->mesh.S;
	by_storage(ctx->storage, [&](auto kind) {
		hipLaunchKernelGGL(cell_release_kernel<decltype(kind)::value>, dim3((S + 63) / 64), dim3(64), 0, ctx->stream,
						   S, (int64_t)index, (float)voxel_volume(ctx), a.vox, ctx->dens, ctx->dscale, ctx->dinv,
						   ctx->rounding_step, a.intern, a.frac);
	});
	BFVM_HIP(hipGetLastError());
}

void launch_voxel_positions(bfvm_ctx* ctx, int count, const float* d_pos, int* d_out)
{
	const cell_geom g = make_geom(ctx);
	hipLaunchKernelGGL(voxel_positions_kernel, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, g, count, d_pos,
					   d_out);
	BFVM_HIP(hipGetLastError());
}

} // namespace bfvm
