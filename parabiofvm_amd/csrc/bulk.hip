// Bulk supply and uptake on gfx950 (host/bulk_solver.cpp:23-35):
//   D = (D + dt*S*T) / (1 + dt*(U + S))        per voxel and substrate.
// The reference asks three per-voxel callbacks for S, U and T at every solve.  Here their values are
// uploaded as [V][S] fields (bfvm_bulk_store_rates), reduced on the host to the numerator term
// num = (dt*S)*T and the denominator den = 1 + dt*(U + S) in the reference's operation order, so one solve
// reads D, num and den and writes D: the FP32 result is the reference's bit for bit.
#include "bfvm_internal.hpp"

namespace bfvm {

// FP32 storage, 4 elements per thread (n a multiple of 4)
__global__ __launch_bounds__(256) void bulk4_kernel(float4* __restrict__ dens, const float4* __restrict__ num,
													const float4* __restrict__ den, int64_t n4)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
	{
		const float4 d = dens[i], a = num[i], b = den[i];
		dens[i] = make_float4((d.x + a.x) / b.x, (d.y + a.y) / b.y, (d.z + a.z) / b.z, (d.w + a.w) / b.w);
	}
}

template <int K>
__global__ __launch_bounds__(256) void bulk_kernel(void* __restrict__ dens, const float* __restrict__ num,
												   const float* __restrict__ den, const float* __restrict__ scale,
												   const float* __restrict__ inv, const unsigned* __restrict__ rounding,
												   int S, int64_t n)
{
	const uint32_t key = K == store_f32 ? 0u : rounding_key(rounding, 8);
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		const int s = (int)(i % S);
		const float d = K == store_f32 ? dload<K, 1>(dens, i) : dload<K, 1>(dens, i) * inv[s];
		const float r = (d + num[i]) / den[i];
		dstore_sr<K, 1>(dens, i, K == store_f32 ? r : r * scale[s], key);
	}
}

void launch_bulk(bfvm_ctx* ctx)
{
	const int64_t n = (int64_t)ctx->mesh.V * ctx->mesh.S;
	if (ctx->storage == BFVM_STORAGE_F32 && n % 4 == 0)
	{
		const int64_t n4 = n / 4;
		const int64_t blocks = std::min<int64_t>(65536, (n4 + 255) / 256);
		hipLaunchKernelGGL(bulk4_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (float4*)ctx->dens,
						   (const float4*)ctx->bulk_num, (const float4*)ctx->bulk_den, n4);
	}
	else
	{
		const int64_t blocks = std::min<int64_t>(65536, (n + 255) / 256);
		by_storage(ctx->storage, [&](auto kind) {
			hipLaunchKernelGGL((bulk_kernel<decltype(kind)::value>), dim3((unsigned)blocks), dim3(256), 0,
							   ctx->stream, ctx->dens, ctx->bulk_num, ctx->bulk_den, ctx->dscale, ctx->dinv,
							   ctx->rounding_step, ctx->mesh.S, n);
		});
	}
	BFVM_HIP(hipGetLastError());
}

} // namespace bfvm
