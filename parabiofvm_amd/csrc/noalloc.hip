// Device scratch for the entry point that allocates nothing (bfvm_voxel_positions_noalloc): a file-scope
// device array holds the staged positions and indices of one call.  Each GPU has its own copy of the array
// at its own address, looked up once per device when the first context is created on it; host code uses that
// address, never the array's name.  The array lives until the process ends, and every call on one GPU uses it:
// two calls that overlap on one GPU (from two host threads) race.
#include <mutex>

#include "bfvm_internal.hpp"

namespace bfvm {

namespace {

constexpr int max_devices = 64;
// points: positions [count][dims <= 3] FP32, then indices [count][3] int32
constexpr size_t positions_bytes = (size_t)noalloc_max_points * 3 * sizeof(float);
constexpr size_t scratch_bytes = positions_bytes + (size_t)noalloc_max_points * 3 * sizeof(int32_t);

__device__ __align__(256) unsigned char scratch[scratch_bytes];

std::mutex g_mutex;
void* g_scratch[max_devices] = {};

} // namespace

void setup_noalloc_scratch(int device)
{
	if (device < 0 || device >= max_devices)
		throw std::invalid_argument("device index beyond the no-allocation scratch table");
	std::lock_guard<std::mutex> lock(g_mutex);
	if (g_scratch[device])
		return;
	void* ptr = nullptr;
	BFVM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
	g_scratch[device] = ptr;
}

void voxel_positions_noalloc(bfvm_ctx* ctx, int count, const float* positions, int32_t* out3)
{
	void* base = nullptr;
	{
		std::lock_guard<std::mutex> lock(g_mutex);
		base = g_scratch[ctx->device];
	}
	if (!base)
		throw std::logic_error("no-allocation scratch not set up on this device");
	float* d_pos = (float*)base;
	int* d_out = (int*)((unsigned char*)base + positions_bytes);
	const size_t dims = (size_t)ctx->mesh.dims;
	BFVM_HIP(hipMemcpyAsync(d_pos, positions, sizeof(float) * count * dims, hipMemcpyHostToDevice, ctx->stream));
	launch_voxel_positions(ctx, count, d_pos, d_out);
	BFVM_HIP(hipMemcpyAsync(out3, d_out, sizeof(int) * count * 3, hipMemcpyDeviceToHost, ctx->stream));
	BFVM_HIP(hipStreamSynchronize(ctx->stream));
}

} // namespace bfvm
