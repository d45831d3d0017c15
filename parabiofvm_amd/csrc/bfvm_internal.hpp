// Internal declarations shared by the HIP translation units of libparabiofvm_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
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

struct graph_slot
{
	hipGraphExec_t exec = nullptr;
	hipGraph_t graph = nullptr;
	uint64_t epoch = 0;
};

} // namespace bfvm

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

	float* dens = nullptr;
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

	bfvm::agents_buffers ag;
	int* ballots = nullptr;
	int* conflict = nullptr; // device flag

	uint64_t epoch = 1; // bumps whenever captured kernel arguments would change
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

void launch_cell(bfvm_ctx* ctx, bool recompute);
void launch_cell_voxels(bfvm_ctx* ctx);
void launch_cell_release(bfvm_ctx* ctx, int index);
void launch_voxel_positions(bfvm_ctx* ctx, int count, const float* d_pos, int* d_out);

int supported_chunk(int request);

} // namespace bfvm
