// Internal declarations shared by the HIP translation units of libparabiofvm_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

#include "parabiofvm_hip.h"

namespace bfvm {

// Vectors of V floats for kernels that move V consecutive substrates per thread (16-byte accesses for V = 4).
template <int V>
struct vec_of
{
	using type = float __attribute__((ext_vector_type(V)));
};
template <>
struct vec_of<1>
{
	using type = float;
};

// Accesses through the global address space.  Device buffers all live in global memory, but pointers read
// from memory (the per-instance arguments of the batched kernels) are generic to the compiler, which then
// emits flat accesses (counted against the LDS wait counter too); loads and stores of HBM data go through
// these helpers so they are global accesses wherever the pointer came from.
template <typename T>
__device__ __forceinline__ T gload(const T* p)
{
	return *(const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, T v)
{
	*(__attribute__((address_space(1))) T*)p = v;
}

template <int V>
__device__ __forceinline__ typename vec_of<V>::type vload(const float* p)
{
	return *(const typename vec_of<V>::type*)p;
}
template <int V>
__device__ __forceinline__ void vstore(float* p, typename vec_of<V>::type v)
{
	*(typename vec_of<V>::type*)p = v;
}
__device__ __forceinline__ float comp(float v, int) { return v; }
__device__ __forceinline__ void set_comp(float& v, int, float x) { v = x; }
template <typename T>
__device__ __forceinline__ float comp(const T& v, int c)
{
	return v[c];
}
template <typename T>
__device__ __forceinline__ void set_comp(T& v, int c, float x)
{
	v[c] = x;
}

// Density storage in HBM: FP32, or BF16 / FP16 holding value * 2^k_s (a per-substrate power-of-two scale).
// Arithmetic is always FP32 in registers.  Conversions to 16 bits round to nearest even (dstore) for values
// that are set (uploads, Dirichlet) and stochastically (dstore_sr) for values that accumulate updates.
enum storage_kind
{
	store_f32 = BFVM_STORAGE_F32,
	store_bf16 = BFVM_STORAGE_BF16,
	store_f16 = BFVM_STORAGE_F16
};

template <int K>
struct store_of
{
	using elem = float;
};
template <>
struct store_of<store_bf16>
{
	using elem = __bf16;
};
template <>
struct store_of<store_f16>
{
	using elem = _Float16;
};

template <int K, int V>
struct svec_of
{
	using type = typename store_of<K>::elem __attribute__((ext_vector_type(V)));
};
template <int K>
struct svec_of<K, 1>
{
	using type = typename store_of<K>::elem;
};

// V consecutive stored densities starting at element `idx`, as FP32 (still scaled)
template <int K, int V>
__device__ __forceinline__ typename vec_of<V>::type dload(const void* base, int64_t idx)
{
	using st = typename svec_of<K, V>::type;
	const st v = gload((const st*)((const typename store_of<K>::elem*)base + idx));
	if constexpr (K == store_f32)
		return v;
	else if constexpr (V == 1)
		return (float)v;
	else
		return __builtin_convertvector(v, typename vec_of<V>::type);
}

// The same load in two halves: the raw stored elements, and their conversion to FP32 (for loads issued one
// loop iteration ahead of their use)
template <int K, int V>
__device__ __forceinline__ typename svec_of<K, V>::type draw(const void* base, int64_t idx)
{
	using st = typename svec_of<K, V>::type;
	return gload((const st*)((const typename store_of<K>::elem*)base + idx));
}
template <int K, int V>
__device__ __forceinline__ typename vec_of<V>::type dcvt(typename svec_of<K, V>::type v)
{
	if constexpr (K == store_f32)
		return v;
	else if constexpr (V == 1)
		return (float)v;
	else
		return __builtin_convertvector(v, typename vec_of<V>::type);
}

template <int K, int V>
__device__ __forceinline__ void dstore(void* base, int64_t idx, typename vec_of<V>::type x)
{
	using st = typename svec_of<K, V>::type;
	st* p = (st*)((typename store_of<K>::elem*)base + idx);
	if constexpr (K == store_f32)
		gstore(p, x);
	else if constexpr (V == 1)
		gstore(p, (typename store_of<K>::elem)x);
	else
		gstore(p, __builtin_convertvector(x, st));
}

// 32-bit integer hash (xorshift-multiply avalanche, "lowbias32" constants of C. Wellons' hash-prospector).
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x)
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

// Per-launch key of the stochastic rounding: the device step counter and a salt naming the store site.
__device__ __forceinline__ uint32_t rounding_key(const unsigned* counter, uint32_t salt)
{
	return mix32(*counter * 16u + salt);
}

// Store of values that accumulate small updates step after step (sweeps, secretion): 16-bit forms round
// stochastically — a uniform random integer below one unit of the dropped bits is added to the FP32 bit
// pattern and the dropped bits are cleared, so the rounding is unbiased and an update smaller than half a
// 16-bit ulp is kept in expectation instead of being lost every step.  BF16 drops 16 bits; FP16 (values
// pre-scaled into its normal range) drops 13.  The truncated value is exactly representable, so the final
// conversion is exact.  Random bits: one hash of the element index with `key` per vector, split into 16-bit
// halves, and a multiplicative-congruential step (an odd multiplier, a bijection of the 32-bit state) for
// every further pair of components.  FP32 stores unchanged.
//
// gfx950 does this conversion in one instruction (v_cvt_sr_bf16_f32 / v_cvt_sr_f16_f32): it adds the top
// 16 (BF16) or 13 (FP16) bits of its random operand below the kept mantissa and truncates, which is the
// scheme above with the random bits taken from the top of the word (measured bit-exact against the
// software form on 2^20 random values).  Component c uses h for even c and h << 16 for odd c.

// Two values converted with stochastic rounding into one pair of 16-bit elements: `a` with random word h,
// `b` with h << 16.
template <int K>
__device__ __forceinline__ typename store_of<K>::elem __attribute__((ext_vector_type(2)))
sr_pair(float a, float b, uint32_t h)
{
	using pair = typename store_of<K>::elem __attribute__((ext_vector_type(2)));
	pair o = {};
	if constexpr (K == store_bf16)
	{
		o = __builtin_amdgcn_cvt_sr_bf16_f32(o, a, h, false);
		o = __builtin_amdgcn_cvt_sr_bf16_f32(o, b, h << 16, true);
	}
	else
	{
		o = __builtin_amdgcn_cvt_sr_f16_f32(o, a, h, false);
		o = __builtin_amdgcn_cvt_sr_f16_f32(o, b, h << 16, true);
	}
	return o;
}

template <int K, int V>
__device__ __forceinline__ void dstore_sr(void* base, int64_t idx, typename vec_of<V>::type x, uint32_t key)
{
	if constexpr (K == store_f32)
		dstore<K, V>(base, idx, x);
	else
	{
		using elem = typename store_of<K>::elem;
		elem* p = (elem*)base + idx;
		uint32_t h = mix32((uint32_t)idx ^ ((uint32_t)((uint64_t)idx >> 32) * 0x9E3779B9u) ^ key);
		if constexpr (V == 1)
			gstore(p, sr_pair<K>(x, 0.f, h)[0]);
		else
		{
			typename svec_of<K, V>::type y;
#pragma unroll
			for (int c = 0; c < V; c += 2)
			{
				if (c > 0)
					h = h * 0x2C1B3C6Du + 0x297A2D39u;
				const auto o = sr_pair<K>(comp(x, c), comp(x, c + 1), h);
				y[c] = o[0];
				y[c + 1] = o[1];
			}
			gstore((typename svec_of<K, V>::type*)p, y);
		}
	}
}

// Stochastic-rounding stores from a per-thread random stream, for code that stores a run of elements in a
// fixed order (a sweep thread's register block): sr_seed hashes the run's first element index with the key
// once, and every further pair of components takes the next word of a xorshift32 generator (Marsaglia 2003;
// shifts 13, 17, 5) — three shift-xor pairs instead of a hash with 32-bit multiplies per element.  The state
// is kept nonzero (zero is the generator's fixed point).
__device__ __forceinline__ uint32_t xorshift32(uint32_t x)
{
	x ^= x << 13;
	x ^= x >> 17;
	x ^= x << 5;
	return x;
}

__device__ __forceinline__ uint32_t sr_seed(int64_t idx, uint32_t key)
{
	return mix32((uint32_t)idx ^ ((uint32_t)((uint64_t)idx >> 32) * 0x9E3779B9u) ^ key) | 1u;
}

template <int K, int V>
__device__ __forceinline__ void dstore_sr_stream(void* base, int64_t idx, typename vec_of<V>::type x, uint32_t& h)
{
	if constexpr (K == store_f32)
		dstore<K, V>(base, idx, x);
	else
	{
		using elem = typename store_of<K>::elem;
		elem* p = (elem*)base + idx;
		if constexpr (V == 1)
		{
			h = xorshift32(h);
			gstore(p, sr_pair<K>(x, 0.f, h)[0]);
		}
		else
		{
			typename svec_of<K, V>::type y;
#pragma unroll
			for (int c = 0; c < V; c += 2)
			{
				h = xorshift32(h);
				const auto o = sr_pair<K>(comp(x, c), comp(x, c + 1), h);
				y[c] = o[0];
				y[c + 1] = o[1];
			}
			gstore((typename svec_of<K, V>::type*)p, y);
		}
	}
}

// Calls f(std::integral_constant<int, K>) for the storage form K of a context.
template <typename F>
inline void by_storage(int storage, F&& f)
{
	if (storage == BFVM_STORAGE_BF16)
		f(std::integral_constant<int, store_bf16>());
	else if (storage == BFVM_STORAGE_F16)
		f(std::integral_constant<int, store_f16>());
	else
		f(std::integral_constant<int, store_f32>());
}

inline size_t storage_bytes(int storage) { return storage == BFVM_STORAGE_F32 ? 4 : 2; }


void set_error(const std::string& msg);

struct hip_failure
{
	hipError_t err;
	const char* what;
	const char* file;
	int line;
};

#define BFVM_HIP(call)                                                                                               \
	do                                                                                                               \
	{                                                                                                                \
		hipError_t err_ = (call);                                                                                    \
		if (err_ != hipSuccess)                                                                                      \
			throw ::bfvm::hip_failure { err_, #call, __FILE__, __LINE__ };                                           \
	} while (0)

// Partition of one axis for the partitioned (modified) Thomas sweep: P blocks, block k has
// base + (k < rem) elements and starts at k*base + min(k, rem).  Block types share coefficients:
//   P == 1: type 0 is the whole line;
//   P >= 2: type 0 = first block, 1 = interior block of base+1, 2 = interior block of base, 3 = last block.
struct partition
{
	int n = 0, P = 0, base = 0, rem = 0;
	int ntypes = 0;
	int mb = 0; // longest block (row stride of the coefficient table)
	int chunk = 0; // register capacity of the kernel instance that runs it
	int lpg = 0;   // lines per workgroup iteration
	int threads = 0;
	int vec = 1; // substrates per thread (1 or 4)
};

// Device-side tables of one axis.
struct axis_tables
{
	int n = 0;
	partition part;
	// exact (reference-order FP32) Thomas coefficients: b' [n][S], c [S], e [n-1][S]
	float* exact_b = nullptr;
	float* exact_c = nullptr;
	float* exact_e = nullptr;
	// partitioned: coef [4][ntypes][mb][S] (r, gamma, alpha' (rho at j=0), gamma'), red [3][2P][S], A [S]
	float* coef = nullptr;
	float* red = nullptr;
	float* A = nullptr;
};

struct agents_buffers
{
	int32_t count = 0;
	size_t capacity = 0;
	float *sec = nullptr, *sat = nullptr, *upt = nullptr, *nex = nullptr, *intern = nullptr, *frac = nullptr;
	float *vol = nullptr, *pos = nullptr;
	float *num = nullptr, *den = nullptr, *fac = nullptr;
	float *rnum = nullptr, *rden = nullptr, *rfac = nullptr;
	int64_t* vox = nullptr; // voxel of each agent (-1 outside the mesh), computed at upload
	int* win = nullptr;		// ballot winner whose slot collects each agent's terms
	// sorted reduction: agents ordered by voxel at upload (keys/idx hold the sort's input and output halves)
	bool sorted = false;
	uint64_t* keys = nullptr; // [2][capacity]
	int* idx = nullptr;		  // [2][capacity]
	float *snum = nullptr, *sden = nullptr, *sfac = nullptr; // the recompute terms in voxel order [capacity][S]
	int* multi = nullptr; // set at upload when some voxel holds more than one agent
	void* sort_tmp = nullptr;
	size_t sort_tmp_bytes = 0;
};

struct mesh_info
{
	int dims = 0;
	int mins[3] = { 0, 0, 0 };
	int maxs[3] = { 0, 0, 0 };
	int vs[3] = { 1, 1, 1 };
	int grid[3] = { 1, 1, 1 };
	int S = 0;
	size_t V = 0;
	float dt = 0;
};

// Work-list item of the batched (many microenvironment) kernels: a range of workgroup-sized units
// (sweep line groups or gradient tiles) of one instance.
struct batch_item
{
	int32_t instance;
	int32_t pad;
	int64_t begin, end;
};

// Per-instance arguments of the batched secretion kernels (one instance's agents and buffers).
struct cell_args
{
	int A, S, with_internalized, voxel_volume;
	float dt;
	const int64_t* vox;
	const float *vol, *sec, *sat, *upt, *nex;
	float *num, *den, *fac, *rnum, *rden, *rfac, *intern;
	int *ballots, *win, *conflict;
	void* dens;
	const float *scale, *inv;
	const unsigned* rounding;
};

constexpr int batch_chunk = 16; // register block of the staged and batched sweeps (one line per thread)
constexpr int batch_max_threads = 1024;

struct graph_slot
{
	hipGraphExec_t exec = nullptr;
	hipGraph_t graph = nullptr;
	uint64_t epoch = 0;
};

} // namespace bfvm

struct bfvm_batch;

struct bfvm_ctx
{
	int device = 0;
	hipStream_t stream = nullptr;
	bool initialized = false;
	bfvm::mesh_info mesh;
	std::vector<float> D, decay;
	float* d_D = nullptr; // device copies [S]
	float* d_decay = nullptr;
	bool compute_internalized = false;

	// densities [V*S] in the storage form (float, __bf16 or _Float16); 16-bit forms hold value * dscale[s]
	void* dens = nullptr;
	int storage = BFVM_STORAGE_F32;		 // form of `dens`
	int storage_request = BFVM_STORAGE_F32; // form the next bfvm_initialize allocates
	float* dscale = nullptr;				 // [S] powers of two (1 for FP32)
	float* dinv = nullptr;					 // [S] 1 / dscale
	std::vector<float> h_scale;
	std::vector<float> bound_max; // [S] largest |Dirichlet value| with its condition set
	// FP16 headroom: value_bound[s] bounds |density| of substrate s (unscaled).  Diffusion with decay keeps
	// values within the range of the previous values and the Dirichlet values; secretion and bulk move a value
	// towards 0 and the agents' saturation / bulk target densities; only net export (export_step per secretion
	// call, all agents' |N| dt / vv summed as if they shared one voxel) and release at death add to it.  When
	// value_bound * scale nears FP16's range the stored values and the scale are halved as often as needed.
	std::vector<double> value_bound, export_step;
	std::vector<float> agent_target, bulk_target; // [S] largest saturation density / bulk target density
	float* rescale_factor = nullptr;			   // [S] device factors of a pending rescale
	unsigned* maxbits = nullptr;  // [S] device scratch of the FP16 scale choice
	unsigned* rounding_step = nullptr; // device counter keying the stochastic rounding of 16-bit stores
	float* bulk_num = nullptr; // [V*S] dt*S*T of the bulk supply/uptake fields (null until uploaded)
	float* bulk_den = nullptr; // [V*S] 1 + dt*(U + S)
	float* staging = nullptr;	  // FP32 staging for host <-> 16-bit conversions
	size_t staging_elems = 0;
	float* grad = nullptr;
	size_t grad_elems = 0;

	// Dirichlet faces: [6][S] values and conditions (face f = 2*axis + is_max); has_faces if any condition set
	float* face_val = nullptr;
	uint8_t* face_cond = nullptr;
	bool has_faces = false;
	// interior Dirichlet, de-duplicated per voxel: linear voxel index, [count][S] values/conditions
	int64_t n_interior = 0;
	int64_t* int_vox = nullptr;
	float* int_val = nullptr;
	uint8_t* int_cond = nullptr;

	bfvm::axis_tables axes[3];
	int algorithm = BFVM_SWEEP_PARTITIONED;
	int chunk_request[3] = { 0, 0, 0 };
	int lines_request[3] = { 0, 0, 0 };
	int gradient_groups = 0; // workgroups the 3-D gradient aims for (0 = automatic)
	bool gradient_stacked = false; // 3-D gradient tiles with the workgroup's waves stacked along y
	int gradient_order = 0; // 3-D gradient workgroup order: 0 = z chunks of a row band per XCD, 1 = plane order
	int sweep_workgroups[3] = { 0, 0, 0 }; // most workgroups of a sweep launch per axis (0 = the built-in cap)

	bfvm::agents_buffers ag;
	int reduction = BFVM_REDUCTION_ATOMIC; // combining of same-voxel agents, applied at the next agents upload
	int* ballots = nullptr;
	int* conflict = nullptr; // device flag

	uint64_t epoch = 1;		 // bumps whenever captured kernel arguments would change
	uint64_t generation = 0; // bumps at every bfvm_initialize (buffers reallocated)
	bfvm::graph_slot graphs[4];

	// batches this context belongs to: a context call first waits for their streams, and a batch step first
	// waits for the context's stream when a context call may have enqueued work on it (stream_used)
	std::vector<bfvm_batch*> batches;
	bool stream_used = false;
};

// Many microenvironments advanced together (bfvm_batch_*): one launch per sweep axis and one for the
// gradient over all instances, driven by device work lists; per-instance kernels for the rest.
struct bfvm_batch
{
	int device = 0;
	hipStream_t stream = nullptr;
	std::vector<bfvm_ctx*> ctx;
	std::vector<uint64_t> generation; // each instance's initialization when the batch was made
	int storage = BFVM_STORAGE_F32;
	unsigned* rounding = nullptr; // step counter of the stochastic rounding, shared by all instances
	struct sweep_plan
	{
		// instances grouped by the vector width of their threads (4, 1), one launch per group
		struct part
		{
			void* inst = nullptr; // device part_args per instance
			bfvm::batch_item* items = nullptr;
			int nitems = 0;
			size_t lds = 0;
			int threads = 0;
			int blocks = 0;
		} parts[2];
		std::vector<float*> tables; // coefficient tables owned by the batch
		std::vector<int> single;	// instances whose mesh has length 1 on this axis (per-instance launch)
		bool any = false;
	} axes[3];
	struct gradient_plan
	{
		void* inst = nullptr; // device gradient_args per instance
		// work items per thread layout: one element, 4 substrates of a voxel, and all of them
		bfvm::batch_item* items[3] = { nullptr, nullptr, nullptr };
		int nitems[3] = { 0, 0, 0 };
		int blocks[3] = { 0, 0, 0 };
		std::vector<int> single; // instances of fewer than 3 dimensions (per-instance launch)
	} grad;
	// batched secretion: per-instance arguments and work items over (agent, substrate) and agent blocks, rebuilt
	// when an instance's agents change (cell_epoch); per-instance launches when an instance uses the sorted
	// reduction
	struct cell_plan
	{
		bfvm::cell_args* inst = nullptr;
		bfvm::batch_item* pair_items = nullptr;
		bfvm::batch_item* agent_items = nullptr;
		size_t pair_cap = 0, agent_cap = 0;
		int npair = 0, nagent = 0;
		bool batched = false;
		uint64_t epoch = ~0ull;
	} cell;
	// launch shapes of the work-list kernels (bfvm_batch_autotune picks them): per axis and thread layout
	// (4-element vectors, one element) the threads a line group aims for and the workgroup cap; for the
	// gradient the y rows per thread, the z planes per tile, the tiles per work item and the workgroup cap
	struct tuning_t
	{
		int group_threads[3][2] = { { 256, 256 }, { 256, 256 }, { 256, 256 } };
		int max_blocks[3][2] = { { 2048, 2048 }, { 2048, 2048 }, { 2048, 2048 } };
		int gradient_tiles = 4;
		int gradient_blocks = 8192;
		int gradient_rows = 4;	  // y rows per thread (4 or 8)
		int gradient_zchunk = 16; // z planes per gradient tile
		int gradient_split = 0;	  // FP32: one gradient launch per thread layout instead of one for both
		int gradient_stacked = 0; // tiles with the workgroup's waves stacked along y
	} tuning;
	uint64_t epoch = 0;
	bfvm::graph_slot graphs[4];
};

namespace bfvm {

// launches (each enqueues on ctx->stream; throw hip_failure on launch errors)
void launch_exact_coefficients(bfvm_ctx* ctx, int axis);
void build_partition(bfvm_ctx* ctx, int axis, int chunk, int lines_request);
void launch_partition_coefficients(bfvm_ctx* ctx, int axis);
void launch_sweep(bfvm_ctx* ctx, int axis);

void launch_dirichlet_faces(bfvm_ctx* ctx);
void launch_dirichlet_interior(bfvm_ctx* ctx);

void launch_gradient(bfvm_ctx* ctx);
void launch_copy(bfvm_ctx* ctx);
// FP32 <-> storage conversion of n densities starting at element `offset` (with the per-substrate scale)
void launch_pack(bfvm_ctx* ctx, const float* src, int64_t offset, int64_t n);
void launch_unpack(bfvm_ctx* ctx, float* dst, int64_t offset, int64_t n);
void launch_max_abs(bfvm_ctx* ctx, const float* src, int64_t n, unsigned* out);
void launch_rescale_f16(bfvm_ctx* ctx, const float* factor, hipStream_t stream);
void launch_advance_rounding(bfvm_ctx* ctx);

void launch_cell(bfvm_ctx* ctx, bool recompute);
void refresh_batch_cell(bfvm_batch* b, uint64_t epoch); // host-synchronous: call outside stream capture
void launch_batched_cell(bfvm_batch* b, bool recompute);
void free_batch_cell(bfvm_batch* b);
void launch_bulk(bfvm_ctx* ctx);
void launch_cell_voxels(bfvm_ctx* ctx);
void launch_cell_sort(bfvm_ctx* ctx); // sorted reduction: order the agents by voxel (upload time)
void launch_cell_release(bfvm_ctx* ctx, int index);
void launch_voxel_positions(bfvm_ctx* ctx, int count, const float* d_pos, int* d_out);

int supported_chunk(int request);
partition choose_partition(const bfvm_ctx* ctx, int axis, int chunk, int lines_request, int max_vec);

// batched kernels (bfvm_batch): plans are built once at bfvm_batch_create; launches enqueue on b->stream
void prepare_batch_sweep(bfvm_batch* b, int axis);
void launch_batched_sweep(bfvm_batch* b, int axis);
void free_batch_sweep(bfvm_batch* b, int axis);
void prepare_batch_gradient(bfvm_batch* b);
void launch_batched_gradient(bfvm_batch* b);
void free_batch_gradient(bfvm_batch* b);

// entry point without allocation: the file-scope device scratch of noalloc.hip, looked up per device
constexpr int noalloc_max_points = 1 << 20;
void setup_noalloc_scratch(int device); // with `device` current, once per device (bfvm_create)
void voxel_positions_noalloc(bfvm_ctx* ctx, int count, const float* positions, int32_t* out3);

// autotuner: time candidate launch shapes on the context's mesh / the batch and keep the fastest
void autotune_context(bfvm_ctx* ctx, bool tune_gradient);
void autotune_batch(bfvm_batch* b);
bool apply_cached_batch_tuning(bfvm_batch* b); // false when this list of instance shapes has no cached tuning
void save_tuning(const char* path);
void load_tuning(const char* path);

} // namespace bfvm
