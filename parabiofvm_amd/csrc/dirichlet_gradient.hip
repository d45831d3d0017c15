// Dirichlet conditions and central-difference gradients on gfx950.
#include <algorithm>

#include "bfvm_internal.hpp"

namespace bfvm {

// ------------------------------------------------------------------------------------------------
// Dirichlet — host/dirichlet_solver.cpp:22-144
// ------------------------------------------------------------------------------------------------

// Boundary faces.  One thread per (face voxel, substrate); every voxel on several faces gets the value of the
// last face in the order x-, x+, y-, y+, z-, z+ whose condition is set (host/dirichlet_solver.cpp:91-110),
// so threads that reach the same edge voxel from different faces write the same value.
// Face voxels are enumerated as the two x planes (2*Y*Z), then the two y planes (2*X*Z), then the two z
// planes (2*X*Y) of the axes the mesh has.
template <int K>
__global__ __launch_bounds__(256) void dirichlet_faces_kernel(void* __restrict__ dens, const float* __restrict__ fval,
															  const uint8_t* __restrict__ fcond,
															  const float* __restrict__ scale, int X, int Y, int Z,
															  int S, int dims, int64_t total)
{
	const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= total)
		return;
	const int s = (int)(idx % S);
	int64_t r = idx / S;
	int x, y, z;
	const int64_t nx = dims >= 1 ? 2 * (int64_t)Y * Z : 0;
	const int64_t ny = dims >= 2 ? 2 * (int64_t)X * Z : 0;
	if (r < nx)
	{
		const int64_t side = r / ((int64_t)Y * Z);
		const int64_t rem = r % ((int64_t)Y * Z);
		x = side ? X - 1 : 0;
		y = (int)(rem % Y);
		z = (int)(rem / Y);
	}
	else if (r < nx + ny)
	{
		r -= nx;
		const int64_t side = r / ((int64_t)X * Z);
		const int64_t rem = r % ((int64_t)X * Z);
		y = side ? Y - 1 : 0;
		x = (int)(rem % X);
		z = (int)(rem / X);
	}
	else
	{
		r -= nx + ny;
		const int64_t side = r / ((int64_t)X * Y);
		const int64_t rem = r % ((int64_t)X * Y);
		z = side ? Z - 1 : 0;
		x = (int)(rem % X);
		y = (int)(rem / X);
	}
	const int coord[3] = { x, y, z };
	const int grid[3] = { X, Y, Z };
	bool set = false;
	float v = 0;
	for (int d = 0; d < dims; d++)
	{
		if (coord[d] == 0 && fcond[(2 * d) * S + s])
		{
			set = true;
			v = fval[(2 * d) * S + s];
		}
		if (coord[d] == grid[d] - 1 && fcond[(2 * d + 1) * S + s])
		{
			set = true;
			v = fval[(2 * d + 1) * S + s];
		}
	}
	if (set)
		dstore<K, 1>(dens, (((int64_t)z * Y + y) * X + x) * S + s, K == store_f32 ? v : v * scale[s]);
}

// Interior list, de-duplicated per voxel at upload (host/dirichlet_solver.cpp:22-44).  One thread per
// (entry, substrate).
template <int K>
__global__ __launch_bounds__(256) void dirichlet_interior_kernel(void* __restrict__ dens,
																 const int64_t* __restrict__ vox,
																 const float* __restrict__ val,
																 const uint8_t* __restrict__ cond,
																 const float* __restrict__ scale, int S,
																 int64_t total)
{
	const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= total)
		return;
	const int64_t i = idx / S;
	const int s = (int)(idx % S);
	if (cond[idx])
		dstore<K, 1>(dens, vox[i] * S + s, K == store_f32 ? val[idx] : val[idx] * scale[s]);
}

void launch_dirichlet_faces(bfvm_ctx* ctx)
{
	if (!ctx->has_faces)
		return;
	const auto& m = ctx->mesh;
	const int X = m.grid[0], Y = m.grid[1], Z = m.grid[2];
	int64_t faces = 2 * (int64_t)Y * Z;
	if (m.dims >= 2)
		faces += 2 * (int64_t)X * Z;
	if (m.dims >= 3)
		faces += 2 * (int64_t)X * Y;
	const int64_t total = faces * m.S;
	by_storage(ctx->storage, [&](auto k) {
		hipLaunchKernelGGL((dirichlet_faces_kernel<decltype(k)::value>), dim3((unsigned)((total + 255) / 256)),
						   dim3(256), 0, ctx->stream, ctx->dens, ctx->face_val, ctx->face_cond, ctx->dscale, X, Y, Z,
						   m.S, m.dims, total);
	});
	BFVM_HIP(hipGetLastError());
}

void launch_dirichlet_interior(bfvm_ctx* ctx)
{
	if (ctx->n_interior == 0)
		return;
	const int64_t total = ctx->n_interior * ctx->mesh.S;
	by_storage(ctx->storage, [&](auto k) {
		hipLaunchKernelGGL((dirichlet_interior_kernel<decltype(k)::value>), dim3((unsigned)((total + 255) / 256)),
						   dim3(256), 0, ctx->stream, ctx->dens, ctx->int_vox, ctx->int_val, ctx->int_cond, ctx->dscale,
						   ctx->mesh.S, total);
	});
	BFVM_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------------
// gradient — host/gradient_solver.cpp:10-291
// ------------------------------------------------------------------------------------------------

// One thread per (voxel, substrate).  Interior: (d[i+1] - d[i-1]) * (1.f/(2 dx)); borders:
// (d[1] - d[0]) * (1.f/dx) and (d[n-1] - d[n-2]) * (1.f/dx), with the float reciprocals of the reference
// (host/gradient_solver.cpp:17-18).  Output layout [d][z][y][x][s] in 3-D, [y][x][s][d] in 1-D/2-D
// (src/traits.h:27-33, 55-62, 86-92).
template <int K>
__global__ __launch_bounds__(256) void gradient_kernel(float* __restrict__ grad, const void* __restrict__ dens,
													   const float* __restrict__ inv, int X, int Y, int Z, int S,
													   int dims, float bx, float cx, float by, float cy, float bz,
													   float cz, int64_t total)
{
	const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= total)
		return;
	const int64_t v = idx / S;
	const int x = (int)(v % X);
	const int y = (int)((v / X) % Y);
	const int z = (int)(v / ((int64_t)X * Y));
	const int64_t sx = S, sy = (int64_t)X * S, sz = (int64_t)X * Y * S;

	const int s = (int)(idx % S);
	const float is = K == store_f32 ? 1.f : inv[s];
	auto p = [&](int64_t o) {
		const float d = dload<K, 1>(dens, idx + o);
		return K == store_f32 ? d : d * is;
	};
	float g[3];
	{
		if (x == 0)
			g[0] = (p(sx) - p(0)) * bx;
		else if (x == X - 1)
			g[0] = (p(0) - p(-sx)) * bx;
		else
			g[0] = (p(sx) - p(-sx)) * cx;
		if (dims >= 2)
		{
			if (y == 0)
				g[1] = (p(sy) - p(0)) * by;
			else if (y == Y - 1)
				g[1] = (p(0) - p(-sy)) * by;
			else
				g[1] = (p(sy) - p(-sy)) * cy;
		}
		if (dims >= 3)
		{
			if (z == 0)
				g[2] = (p(sz) - p(0)) * bz;
			else if (z == Z - 1)
				g[2] = (p(0) - p(-sz)) * bz;
			else
				g[2] = (p(sz) - p(-sz)) * cz;
		}
	}
	if (dims == 3)
	{
		grad[idx] = g[0];
		grad[total + idx] = g[1];
		grad[2 * total + idx] = g[2];
	}
	else
	{
		for (int d = 0; d < dims; d++)
			grad[idx * dims + d] = g[d];
	}
}

// 3-D gradient, one thread per (TY rows of y, q) column group (q = x*S + s, V consecutive substrates of one
// voxel per thread) marching along z with the z-1, z, z+1 densities of its rows in registers.  The x
// neighbours come from the lanes holding voxel x +- 1 (cross-lane reads), the y rows just outside the
// thread's rows are read from the current plane.  Tiles are numbered z chunk innermost; the workgroup order
// over them is a launch choice (gradient3d_kernel).  Index arithmetic is 32-bit per plane; plane offsets are
// 64-bit.
constexpr int grad_tile = 256;
constexpr int flat4_vec = 8; // gradient_args::vec of a one-element instance tiled 4 row positions per thread
constexpr int num_xcds = 8;



// One 3-D gradient field: the densities, the output and the tiling of its (y, q) columns.
struct gradient_args
{
	float* grad;
	const void* dens;
	const float* inv;
	int Y, Z, S, XS, vec;
	int tiles_per_row, zchunks, zchunk, ntiles;
	float bx, cx, by, cy, bz, cz;
	int64_t VS;
};

// Rows of y per thread in the 3-D gradient (TY): the z window of every row stays in registers and the rows
// of a thread share their y neighbours, so one plane step loads TY + 2 rows for TY outputs.  4 rows: on
// 400^3 x 8 FP32 the tuner's best shape runs 1.72 ms with 4 rows (stacked tiles, whose 4 waves share their
// halo rows through LDS) against 1.78 ms with 8 (side-by-side tiles; 8 rows had been 1.89 ms against 2.42 with
// one row before the z-chunked tiles); 16-bit storage and the batched form are slower with 8.
template <int V, int K>
constexpr int gradient_rows()
{
	return 4;
}

constexpr int wave_lanes = 64;

// v as held by lane `src` of this wave (ds_bpermute; src is taken modulo the wave size)
template <int V>
__device__ __forceinline__ typename vec_of<V>::type shfl_vec(typename vec_of<V>::type v, int src)
{
	typename vec_of<V>::type r;
#pragma unroll
	for (int c = 0; c < V; c++)
		set_comp(r, c, __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(comp(v, c)))));
	return r;
}

template <int V, int K, int TY>
__device__ __forceinline__ void gradient_tile(const gradient_args& a, int tile)
{
	using vec = typename vec_of<V>::type;
	// z chunk innermost: the tiles one XCD runs together are consecutive z ranges of one row band, so the
	// planes where one chunk ends and the next begins are read from its L2 (measured 2.5 % faster in FP32
	// and 7 % in BF16 than keeping neighbouring row bands together)
	const int zc = tile % a.zchunks;
	const int row_tile = tile / a.zchunks;
	const int y0 = (row_tile / a.tiles_per_row) * TY;
	const int q = ((row_tile % a.tiles_per_row) * grad_tile + threadIdx.x) * V;
	const int S = a.S, XS = a.XS, Y = a.Y, Z = a.Z;
	if (q >= XS)
		return;
	const int z0 = zc * a.zchunk;
	const int z1 = min(Z, z0 + a.zchunk);
	const int ny = min(TY, Y - y0); // rows of this thread
	const int64_t plane = (int64_t)Y * XS;
	const int64_t off = (int64_t)y0 * XS + q;
	float* g0 = a.grad + off;
	float* g1 = g0 + a.VS;
	float* g2 = g1 + a.VS;
	const bool xlo = q < S, xhi = q >= XS - S;
	const int lane = threadIdx.x & (wave_lanes - 1);
	const int SG = S / V; // lanes per voxel
	const float bx = a.bx, cx = a.cx, by = a.by, cy = a.cy, bz = a.bz, cz = a.cz;
	const void* dens = a.dens;

	vec is = vec(1.f);
	if constexpr (K != store_f32)
		is = vload<V>(a.inv + q % S);
	auto ld = [&](int64_t o) {
		const vec d = dload<K, V>(dens, off + o);
		if constexpr (K == store_f32)
			return d;
		else
			return d * is;
	};
	vec zm[TY], c[TY];
#pragma unroll
	for (int r = 0; r < TY; r++)
	{
		zm[r] = (r < ny && z0 > 0) ? ld((int64_t)(z0 - 1) * plane + (int64_t)r * XS) : vec(0.f);
		c[r] = r < ny ? ld((int64_t)z0 * plane + (int64_t)r * XS) : vec(0.f);
	}
	// 16-bit forms keep one plane step of loads in flight: plane z + 1 and the y neighbours of plane z are
	// loaded (raw) during the step before, as a z step is otherwise bound by the latency of its own loads
	constexpr bool prefetch = K != store_f32;
	using raw = typename svec_of<K, V>::type;
	raw pzp[prefetch ? TY : 1], pbelow, pabove;
	auto fetch = [&](int zz) {
		const int64_t oo = (int64_t)zz * plane;
#pragma unroll
		for (int r = 0; r < TY; r++)
			if (r < ny && zz + 1 < Z)
				pzp[r] = draw<K, V>(dens, off + oo + plane + (int64_t)r * XS);
		if (y0 > 0)
			pbelow = draw<K, V>(dens, off + oo - XS);
		if (y0 + ny < Y)
			pabove = draw<K, V>(dens, off + oo + (int64_t)ny * XS);
	};
	if constexpr (prefetch)
		if (z0 < z1)
			fetch(z0);
	for (int z = z0; z < z1; z++)
	{
		const int64_t o = (int64_t)z * plane;
		vec zp[TY];
		vec below, above;
		if constexpr (prefetch)
		{
#pragma unroll
			for (int r = 0; r < TY; r++)
				zp[r] = (r < ny && z + 1 < Z) ? dcvt<K, V>(pzp[r]) * is : vec(0.f);
			below = y0 > 0 ? dcvt<K, V>(pbelow) * is : vec(0.f);
			above = y0 + ny < Y ? dcvt<K, V>(pabove) * is : vec(0.f);
			if (z + 1 < z1)
				fetch(z + 1);
		}
		else
		{
#pragma unroll
			for (int r = 0; r < TY; r++)
				zp[r] = (r < ny && z + 1 < Z) ? ld(o + plane + (int64_t)r * XS) : vec(0.f);
			// rows y0-1 and y0+ny (the y neighbours outside this thread's rows)
			below = y0 > 0 ? ld(o - XS) : vec(0.f);
			above = y0 + ny < Y ? ld(o + (int64_t)ny * XS) : vec(0.f);
		}
#pragma unroll
		for (int r = 0; r < TY; r++)
		{
			if (r >= ny)
				break;
			const int y = y0 + r;
			const int64_t orow = o + (int64_t)r * XS;
			// x neighbours.  One element per thread: loaded, with the address clamped to the thread's own
			// element at the row's ends (the loads hit the neighbouring lanes' lines in cache; measured on the
			// 64-instance batch, whose one-element instances hold 70 % of its densities: gradient FP32 4.16 ->
			// 3.89 ms, FP16 3.91 -> 3.77 ms, against lane shuffles with edge loads under divergent branches).
			// Vectors: the same substrates of voxel x +- 1 sit SG lanes away in this wave's registers; only
			// the lanes at the wave's edges load theirs.
			vec xp, xm;
			if constexpr (V == 1)
			{
				xp = ld(xhi ? orow : orow + S);
				xm = ld(xlo ? orow : orow - S);
			}
			else if constexpr (K != store_f32)
			{
				// 16-bit: both edge loads issued before either is converted (converted inside its own branch,
				// each load was waited for where it was issued)
				xp = shfl_vec<V>(c[r], lane + SG);
				xm = shfl_vec<V>(c[r], lane - SG);
				const bool ep = !xhi && lane + SG >= wave_lanes, em = !xlo && lane < SG;
				raw rp, rm;
				if (ep)
					rp = draw<K, V>(dens, off + orow + S);
				if (em)
					rm = draw<K, V>(dens, off + orow - S);
				if (xhi)
					xp = vec(0.f);
				else if (ep)
					xp = dcvt<K, V>(rp) * is;
				if (xlo)
					xm = vec(0.f);
				else if (em)
					xm = dcvt<K, V>(rm) * is;
			}
			else
			{
				xp = shfl_vec<V>(c[r], lane + SG);
				xm = shfl_vec<V>(c[r], lane - SG);
				if (xhi)
					xp = vec(0.f);
				else if (lane + SG >= wave_lanes)
					xp = ld(orow + S);
				if (xlo)
					xm = vec(0.f);
				else if (lane < SG)
					xm = ld(orow - S);
			}
			const vec ym = r == 0 ? below : c[r - 1 < 0 ? 0 : r - 1];
			const vec yp = r == ny - 1 ? above : c[r + 1 < TY ? r + 1 : TY - 1];
			vec gx, gy, gz;
			if (xlo)
				gx = (xp - c[r]) * bx;
			else if (xhi)
				gx = (c[r] - xm) * bx;
			else
				gx = (xp - xm) * cx;
			if (y == 0)
				gy = (yp - c[r]) * by;
			else if (y == Y - 1)
				gy = (c[r] - ym) * by;
			else
				gy = (yp - ym) * cy;
			if (z == 0)
				gz = (zp[r] - c[r]) * bz;
			else if (z == Z - 1)
				gz = (c[r] - zm[r]) * bz;
			else
				gz = (zp[r] - zm[r]) * cz;
			// streaming outputs: non-temporal stores keep them from displacing the density rows in L2
			const int64_t go = o + (int64_t)r * XS;
			__builtin_nontemporal_store(gx, (vec*)(g0 + go));
			__builtin_nontemporal_store(gy, (vec*)(g1 + go));
			__builtin_nontemporal_store(gz, (vec*)(g2 + go));
		}
#pragma unroll
		for (int r = 0; r < TY; r++)
		{
			zm[r] = c[r];
			c[r] = zp[r];
		}
	}
}

// One-element instances (S % 4 != 0) of batches in vectors of 4 consecutive row positions q..q+3
// (q = x S + s): each lane moves 16 (FP32) or 8 bytes per access instead of 4 or 2, with a quarter of the
// memory instructions.  Measured on the 64-instance batch (its one-element instances hold 70 % of its
// densities): gradient FP32 3.89 -> 3.09 ms, FP16 3.77 -> 3.28 ms (16-bit with the one-plane prefetch; without
// it 4.11 ms).  The
// positions of a vector belong to different substrates and voxels: rows need not start on a vector boundary,
// so accesses are unaligned vectors; the x neighbours (q +- S) are unaligned vector loads (cache hits: the
// neighbouring lanes' own row); the x boundary is decided per component (x = 0 where q + c < S, x = X - 1 where
// q + c >= XS - S).  A vector that reaches past its row's end moves component by component.
template <int K>
struct uvec
{
	using type = typename store_of<K>::elem __attribute__((ext_vector_type(4), aligned(sizeof(typename store_of<K>::elem))));
};
using float4u = float __attribute__((ext_vector_type(4), aligned(4)));

template <int K, int TY>
__device__ __forceinline__ void gradient_tile_flat4(const gradient_args& a, int tile)
{
	using vec = typename vec_of<4>::type;
	using elem = typename store_of<K>::elem;
	using uv = typename uvec<K>::type;
	const int zc = tile % a.zchunks;
	const int row_tile = tile / a.zchunks;
	const int y0 = (row_tile / a.tiles_per_row) * TY;
	const int q = ((row_tile % a.tiles_per_row) * grad_tile + threadIdx.x) * 4;
	const int S = a.S, XS = a.XS, Y = a.Y, Z = a.Z;
	if (q >= XS)
		return;
	const bool whole = q + 4 <= XS;
	const int z0 = zc * a.zchunk;
	const int z1 = min(Z, z0 + a.zchunk);
	const int ny = min(TY, Y - y0);
	const int64_t plane = (int64_t)Y * XS;
	const int64_t off = (int64_t)y0 * XS + q;
	float* g0 = a.grad + off;
	float* g1 = g0 + a.VS;
	float* g2 = g1 + a.VS;
	const float bx = a.bx, cx = a.cx, by = a.by, cy = a.cy, bz = a.bz, cz = a.cz;
	const elem* dens = (const elem*)a.dens + off;
	vec is = vec(1.f);
	if constexpr (K != store_f32)
	{
#pragma unroll
		for (int c = 0; c < 4; c++)
			is[c] = a.inv[(q + c) % S];
	}
	// Loads take 4 consecutive elements as one vector wherever the 4 lie inside the array (vin), also where
	// they cross the row's start or end: the elements of the neighbouring row feed only components that are not
	// stored (past the row's end) or x neighbours that one-sided differences do not use (x at a boundary).  The
	// per-component form remains for the array's first and last elements.  Tested per lane on the row's
	// bounds (whole, inner), nearly every wave held a lane at a row's end and ran both forms.
	auto vin = [&](int64_t o) {
		const int64_t e = off + o;
		return e >= 0 && e + 4 <= a.VS;
	};
	// 4 positions from element o of this thread's row position; components past the row's end read as 0
	auto ld = [&](int64_t o) -> vec {
		vec r;
		if (vin(o))
			r = __builtin_convertvector(gload((const uv*)(dens + o)), vec);
		else
		{
#pragma unroll
			for (int c = 0; c < 4; c++)
				r[c] = q + c < XS ? (float)gload(dens + o + c) : 0.f;
		}
		if constexpr (K == store_f32)
			return r;
		else
			return r * is;
	};
	// the x neighbours d = +S or -S of the 4 positions; components whose neighbour lies outside the row (x at
	// a boundary, where the one-sided difference does not use it) read their own value
	auto ldx = [&](int64_t o, int d) -> vec {
		vec r;
		if (vin(o + d))
			r = __builtin_convertvector(gload((const uv*)(dens + o + d)), vec);
		else
		{
#pragma unroll
			for (int c = 0; c < 4; c++)
			{
				const int p = q + c + d;
				const bool in = q + c < XS && p >= 0 && p < XS;
				r[c] = in ? (float)gload(dens + o + c + d) : 0.f;
			}
		}
		if constexpr (K == store_f32)
			return r;
		else
			return r * is;
	};
	auto st = [&](float* g, int64_t o, vec v) {
		if (whole)
			__builtin_nontemporal_store((float4u)v, (float4u*)(g + o));
		else
		{
#pragma unroll
			for (int c = 0; c < 4; c++)
				if (q + c < XS)
					__builtin_nontemporal_store(v[c], g + o + c);
		}
	};
	vec zm[TY], cc[TY];
#pragma unroll
	for (int r = 0; r < TY; r++)
	{
		zm[r] = (r < ny && z0 > 0) ? ld((int64_t)(z0 - 1) * plane + (int64_t)r * XS) : vec(0.f);
		cc[r] = r < ny ? ld((int64_t)z0 * plane + (int64_t)r * XS) : vec(0.f);
	}
	// 16-bit: plane z + 1 and the y neighbours of plane z are loaded (raw) one plane step ahead
	constexpr bool prefetch = K != store_f32;
	using raw = typename svec_of<K, 4>::type;
	raw pzp[prefetch ? TY : 1], pbelow, pabove;
	auto rld = [&](int64_t o) -> raw {
		raw r;
		if (vin(o))
			r = gload((const uv*)(dens + o));
		else
		{
#pragma unroll
			for (int c = 0; c < 4; c++)
				r[c] = q + c < XS ? gload(dens + o + c) : (elem)0.f;
		}
		return r;
	};
	auto cvt = [&](raw r) -> vec { return __builtin_convertvector(r, vec) * is; };
	// ldx in two halves for 16-bit forms: both x-neighbour loads of a row are issued before either is
	// converted (converted inside its branch, each load was waited for where it was issued)
	auto rldx = [&](int64_t o, int d) -> raw {
		raw r;
		if (vin(o + d))
			r = gload((const uv*)(dens + o + d));
		else
		{
#pragma unroll
			for (int c = 0; c < 4; c++)
			{
				const int p = q + c + d;
				const bool in = q + c < XS && p >= 0 && p < XS;
				r[c] = in ? gload(dens + o + c + d) : (elem)0.f;
			}
		}
		return r;
	};
	auto fetch = [&](int zz) {
		const int64_t oo = (int64_t)zz * plane;
#pragma unroll
		for (int r = 0; r < TY; r++)
			if (r < ny && zz + 1 < Z)
				pzp[r] = rld(oo + plane + (int64_t)r * XS);
		if (y0 > 0)
			pbelow = rld(oo - XS);
		if (y0 + ny < Y)
			pabove = rld(oo + (int64_t)ny * XS);
	};
	if constexpr (prefetch)
		if (z0 < z1)
			fetch(z0);
	for (int z = z0; z < z1; z++)
	{
		const int64_t o = (int64_t)z * plane;
		vec zp[TY];
		vec below, above;
		if constexpr (prefetch)
		{
#pragma unroll
			for (int r = 0; r < TY; r++)
				zp[r] = (r < ny && z + 1 < Z) ? cvt(pzp[r]) : vec(0.f);
			below = y0 > 0 ? cvt(pbelow) : vec(0.f);
			above = y0 + ny < Y ? cvt(pabove) : vec(0.f);
			if (z + 1 < z1)
				fetch(z + 1);
		}
		else
		{
#pragma unroll
			for (int r = 0; r < TY; r++)
				zp[r] = (r < ny && z + 1 < Z) ? ld(o + plane + (int64_t)r * XS) : vec(0.f);
			below = y0 > 0 ? ld(o - XS) : vec(0.f);
			above = y0 + ny < Y ? ld(o + (int64_t)ny * XS) : vec(0.f);
		}
#pragma unroll
		for (int r = 0; r < TY; r++)
		{
			if (r >= ny)
				break;
			const int y = y0 + r;
			const int64_t orow = o + (int64_t)r * XS;
			vec xp, xm;
			if constexpr (K == store_f32)
			{
				xp = ldx(orow, S);
				xm = ldx(orow, -S);
			}
			else
			{
				const raw rp = rldx(orow, S), rm = rldx(orow, -S);
				xp = cvt(rp);
				xm = cvt(rm);
			}
			const vec ym = r == 0 ? below : cc[r - 1 < 0 ? 0 : r - 1];
			const vec yp = r == ny - 1 ? above : cc[r + 1 < TY ? r + 1 : TY - 1];
			vec gx, gy, gz;
#pragma unroll
			for (int c = 0; c < 4; c++)
			{
				const bool xlo = q + c < S, xhi = q + c >= XS - S;
				gx[c] = xlo ? (xp[c] - cc[r][c]) * bx : (xhi ? (cc[r][c] - xm[c]) * bx : (xp[c] - xm[c]) * cx);
			}
			if (y == 0)
				gy = (yp - cc[r]) * by;
			else if (y == Y - 1)
				gy = (cc[r] - ym) * by;
			else
				gy = (yp - ym) * cy;
			if (z == 0)
				gz = (zp[r] - cc[r]) * bz;
			else if (z == Z - 1)
				gz = (cc[r] - zm[r]) * bz;
			else
				gz = (zp[r] - zm[r]) * cz;
			const int64_t go = orow;
			st(g0, go, gx);
			st(g1, go, gy);
			st(g2, go, gz);
		}
#pragma unroll
		for (int r = 0; r < TY; r++)
		{
			zm[r] = cc[r];
			cc[r] = zp[r];
		}
	}
}


// The same gradient with the four waves of a workgroup stacked along y: each wave takes a 64-column segment
// (q) and TY rows, the waves consecutive TY-row bands of that segment.  The rows where two waves meet are
// passed through LDS (each wave publishes its first and last row of the plane, double-buffered by plane parity,
// one barrier per plane), so a workgroup reads its 4 TY rows plus 2 halo rows per plane instead of TY + 2 per
// TY rows, and partly filled tiles at a row's end idle at most one wave-width of lanes.  Results are the same
// expressions on the same values as gradient_tile, bit for bit.  Every wave runs every barrier, also waves
// whose rows or columns lie outside the mesh.
// Halo lanes (16-bit storage, 4 substrates per lane): the S/4 lanes at each end of a wave load the columns
// next to the wave's segment and store nothing, so every x neighbour comes from a lane of the wave and no
// lane loads one by itself.  400^3 x 8 FP16: 1.45 -> 1.38 ms.  In FP32 the same mapping ran 1.46 -> 1.57 ms
// (60-lane segments start off the 128-byte lines, and FP32 rows are twice as long), so FP32 keeps 64-lane
// segments and loads the two edge columns of each row.  Needs S / 4 < 32 (stacked_gradient_ok).
template <int V, int K, int TY, bool PF = false>
__device__ __forceinline__ void gradient_tile_stacked(const gradient_args& a, int tile, typename vec_of<V>::type* sh)
{
	using vec = typename vec_of<V>::type;
	const int zc = tile % a.zchunks;
	const int row_tile = tile / a.zchunks;
	const int wave = threadIdx.x / wave_lanes;
	const int lane = threadIdx.x & (wave_lanes - 1);
	const int y0 = (row_tile / a.tiles_per_row) * (4 * TY) + wave * TY;
	constexpr bool halo = V == 4 && K != store_f32;
	const int SGh = halo ? a.S / V : 0; // halo lanes at each end of the wave
	const int q = ((row_tile % a.tiles_per_row) * (wave_lanes - 2 * SGh) + lane - SGh) * V;
	const int S = a.S, XS = a.XS, Y = a.Y, Z = a.Z;
	const bool col = q >= 0 && q < XS;
	const int ny = max(0, min(TY, Y - y0)); // rows of this wave (0: outside the mesh, barriers only)
	const bool run = col && ny > 0; // loads
	const bool outp = run && lane >= SGh && lane < wave_lanes - SGh; // computes and stores
	const int z0 = zc * a.zchunk;
	const int z1 = min(Z, z0 + a.zchunk);
	const int64_t plane = (int64_t)Y * XS;
	const int64_t off = (int64_t)max(0, min(y0, Y - 1)) * XS + (col ? q : 0);
	float* g0 = a.grad + off;
	float* g1 = g0 + a.VS;
	float* g2 = g1 + a.VS;
	const bool xlo = q < S, xhi = q >= XS - S;
	const int SG = S / V;
	const float bx = a.bx, cx = a.cx, by = a.by, cy = a.cy, bz = a.bz, cz = a.cz;
	const void* dens = a.dens;
	// neighbours in y: the wave below / above when it has rows, else the plane's halo row in memory
	const bool below_wave = wave > 0;
	const bool above_wave = wave < 3 && y0 + ny < Y && ny == TY;

	vec is = vec(1.f);
	if constexpr (K != store_f32)
		if (col)
			is = vload<V>(a.inv + q % S);
	auto ld = [&](int64_t o) {
		const vec d = dload<K, V>(dens, off + o);
		if constexpr (K == store_f32)
			return d;
		else
			return d * is;
	};
	vec zm[TY], c[TY];
#pragma unroll
	for (int r = 0; r < TY; r++)
	{
		zm[r] = (run && r < ny && z0 > 0) ? ld((int64_t)(z0 - 1) * plane + (int64_t)r * XS) : vec(0.f);
		c[r] = (run && r < ny) ? ld((int64_t)z0 * plane + (int64_t)r * XS) : vec(0.f);
	}
	// PF (16-bit forms of the single-instance kernel): load plane z + 1 and the halo rows of plane z (raw) one
	// plane step ahead, as the side-by-side tile does.  Loaded and converted in the same step, each conditional
	// 16-bit load was waited for where it was issued (the conversion sits inside its branch), so the step's
	// loads ran one after another: 400^3 x 8 FP16 stacked tiles 2.05-2.42 ms, 1.45 ms with PF.  The batched
	// kernel keeps the form without: with PF (3 waves per SIMD instead of 4) the batch autotuner picked stacked
	// tiles that ran slower inside the step (64-instance FP16 gradient 3.28 -> 3.50 ms).
	constexpr bool prefetch = PF && K != store_f32;
	using raw = typename svec_of<K, V>::type;
	raw pzp[prefetch ? TY : 1], pbelow, pabove;
	const bool need_below = outp && !below_wave && y0 > 0;
	const bool need_above = outp && !above_wave && y0 + ny < Y;
	auto fetch = [&](int zz) {
		const int64_t oo = (int64_t)zz * plane;
#pragma unroll
		for (int r = 0; r < TY; r++)
			if (run && r < ny && zz + 1 < Z)
				pzp[r] = draw<K, V>(dens, off + oo + plane + (int64_t)r * XS);
		if (need_below)
			pbelow = draw<K, V>(dens, off + oo - XS);
		if (need_above)
			pabove = draw<K, V>(dens, off + oo + (int64_t)ny * XS);
	};
	if constexpr (prefetch)
		if (z0 < z1)
			fetch(z0);
	for (int z = z0; z < z1; z++)
	{
		const int64_t o = (int64_t)z * plane;
		// publish this wave's first and last row of plane z
		vec* buf = sh + (z & 1) * (2 * 4 * wave_lanes);
		buf[(2 * wave) * wave_lanes + lane] = c[0];
		buf[(2 * wave + 1) * wave_lanes + lane] = c[ny > 0 ? ny - 1 : 0];
		vec zp[TY];
		vec below = vec(0.f), above = vec(0.f);
		if constexpr (prefetch)
		{
#pragma unroll
			for (int r = 0; r < TY; r++)
				zp[r] = (run && r < ny && z + 1 < Z) ? dcvt<K, V>(pzp[r]) * is : vec(0.f);
			if (need_below)
				below = dcvt<K, V>(pbelow) * is;
			if (need_above)
				above = dcvt<K, V>(pabove) * is;
			if (z + 1 < z1)
				fetch(z + 1);
		}
		else
		{
#pragma unroll
			for (int r = 0; r < TY; r++)
				zp[r] = (run && r < ny && z + 1 < Z) ? ld(o + plane + (int64_t)r * XS) : vec(0.f);
			if (need_below)
				below = ld(o - XS);
			if (need_above)
				above = ld(o + (int64_t)ny * XS);
		}
		__syncthreads();
		if (run)
		{
			if (below_wave)
				below = buf[(2 * (wave - 1) + 1) * wave_lanes + lane];
			if (above_wave)
				above = buf[(2 * (wave + 1)) * wave_lanes + lane];
		}
		// the edge lanes' x neighbours of every row of the plane, loaded before the row loop: loaded inside it,
		// each row waited for its two loads and with them for the stores of the rows before (400^3 x 8 FP32
		// 1.46 -> 1.42 ms, although the registers they hold cost a wave per SIMD)
		constexpr bool hoist = !halo && !prefetch;
		vec xpe[hoist ? TY : 1], xme[hoist ? TY : 1];
		if constexpr (hoist)
		{
			const bool ep = outp && !xhi && lane + SG >= wave_lanes, em = outp && !xlo && lane < SG;
#pragma unroll
			for (int r = 0; r < TY; r++)
			{
				if (ep && r < ny)
					xpe[r] = ld(o + (int64_t)r * XS + S);
				if (em && r < ny)
					xme[r] = ld(o + (int64_t)r * XS - S);
			}
		}
#pragma unroll
		for (int r = 0; r < TY; r++)
		{
			// x neighbours across lanes (every lane of the wave takes part in the exchange)
			vec xp = shfl_vec<V>(c[r], lane + SG);
			vec xm = shfl_vec<V>(c[r], lane - SG);
			if (!outp || r >= ny)
				continue;
			const int y = y0 + r;
			const int64_t orow = o + (int64_t)r * XS;
			if constexpr (halo)
			{
				if (xhi)
					xp = vec(0.f);
				if (xlo)
					xm = vec(0.f);
			}
			else if constexpr (prefetch)
			{
				// both edge loads issued before either is converted (one wait instead of two)
				const bool ep = !xhi && lane + SG >= wave_lanes, em = !xlo && lane < SG;
				raw rp, rm;
				if (ep)
					rp = draw<K, V>(dens, off + orow + S);
				if (em)
					rm = draw<K, V>(dens, off + orow - S);
				if (xhi)
					xp = vec(0.f);
				else if (ep)
					xp = dcvt<K, V>(rp) * is;
				if (xlo)
					xm = vec(0.f);
				else if (em)
					xm = dcvt<K, V>(rm) * is;
			}
			else
			{
				if (xhi)
					xp = vec(0.f);
				else if (lane + SG >= wave_lanes)
					xp = xpe[r];
				if (xlo)
					xm = vec(0.f);
				else if (lane < SG)
					xm = xme[r];
			}
			const vec ym = r == 0 ? below : c[r - 1 < 0 ? 0 : r - 1];
			const vec yp = r == ny - 1 ? above : c[r + 1 < TY ? r + 1 : TY - 1];
			vec gx, gy, gz;
			if (xlo)
				gx = (xp - c[r]) * bx;
			else if (xhi)
				gx = (c[r] - xm) * bx;
			else
				gx = (xp - xm) * cx;
			if (y == 0)
				gy = (yp - c[r]) * by;
			else if (y == Y - 1)
				gy = (c[r] - ym) * by;
			else
				gy = (yp - ym) * cy;
			if (z == 0)
				gz = (zp[r] - c[r]) * bz;
			else if (z == Z - 1)
				gz = (c[r] - zm[r]) * bz;
			else
				gz = (zp[r] - zm[r]) * cz;
			const int64_t go = o + (int64_t)r * XS;
			__builtin_nontemporal_store(gx, (vec*)(g0 + go));
			__builtin_nontemporal_store(gy, (vec*)(g1 + go));
			__builtin_nontemporal_store(gz, (vec*)(g2 + go));
		}
#pragma unroll
		for (int r = 0; r < TY; r++)
		{
			zm[r] = c[r];
			c[r] = zp[r];
		}
	}
	__syncthreads(); // the next tile of this workgroup writes the same LDS rows
}

// Workgroup order over the tiles (tile = row tile * zchunks + z chunk):
//  - order 0: consecutive tiles on one XCD, which the dispatcher hands workgroup b % 8, so workgroup b takes
//    tile (b % 8) * per_xcd + b / 8: one XCD runs consecutive z chunks of a row band, and the planes where one
//    chunk ends and the next begins come from its L2;
//  - order 1: plane order, workgroup b takes row tile b % row_tiles of z chunk b / row_tiles, so the
//    workgroups resident together write a narrow window of each output (one z chunk of every row band).
// Measured on 400^3 x 8: FP32 1.72 ms (order 0) against 1.46 ms (order 1); FP16 storage 1.82 against 1.95 ms.
// A probe of the gradient's traffic alone (read one array, write three, marching along z) showed the same:
// 8-plane chunks ran 1.49 ms in plane order and 1.59-1.63 ms with each XCD, or each workgroup wave, spread
// over many z chunks.  The autotuner picks the order.
template <int V, int K, bool STACKED>
__global__ __launch_bounds__(grad_tile) void gradient3d_kernel(gradient_args a, int per_xcd, int order)
{
	const int b = blockIdx.x;
	int tile;
	if (order == 0)
		tile = (b % num_xcds) * per_xcd + b / num_xcds;
	else
	{
		const int row_tiles = a.ntiles / a.zchunks;
		tile = b < a.ntiles ? (b % row_tiles) * a.zchunks + b / row_tiles : a.ntiles;
	}
	if (tile >= a.ntiles)
		return;
	if constexpr (STACKED)
	{
		__shared__ typename vec_of<V>::type sh[2 * 2 * 4 * wave_lanes];
		gradient_tile_stacked<V, K, gradient_rows<V, K>(), true>(a, tile, sh);
	}
	else
		gradient_tile<V, K, gradient_rows<V, K>()>(a, tile);
}

// Many microenvironments: workgroups walk (instance, tile range) items, 4 substrates per thread for the
// instances whose substrate count 4 divides, else one element.  V = 0 takes the layout of each instance at run
// time (one launch for the whole batch); V = 4 or 1 serves one layout, so that layout gets its own register
// budget.  Measured on the 64-instance batch: FP32 4.38 ms with one launch against 4.46 ms split by layout,
// 16-bit 4.25 against 4.14 ms (the 16-bit one-element instances need fewer registers than the 4-wide ones).
// TY: y rows per thread (4 or 8, a tuning choice: 8 rows share their y neighbours over more outputs but hold
// more registers).
template <int K, int V, int TY, bool STACKED>
__global__ __launch_bounds__(grad_tile) void gradient_batched_kernel(const gradient_args* __restrict__ inst,
																	 const batch_item* __restrict__ items, int nitems)
{
	using vec4 = typename vec_of<4>::type;
	__shared__ vec4 sh[STACKED ? 2 * 2 * 4 * wave_lanes : 1];
	for (int it = blockIdx.x; it < nitems; it += gridDim.x)
	{
		const batch_item w = items[it];
		const gradient_args a = inst[w.instance];
		for (int64_t t = w.begin; t < w.end; t++)
		{
			if constexpr (STACKED)
			{
				if (V == 4 || (V == 0 && a.vec == 4))
					gradient_tile_stacked<4, K, TY>(a, (int)t, sh);
				else
					gradient_tile_stacked<1, K, TY>(a, (int)t, (float*)sh);
			}
			else if constexpr (V == 4)
				gradient_tile<4, K, TY>(a, (int)t);
			else
			{
				// layouts decided at run time; each launch holds only the tile forms it can meet, so the
				// one-element launch keeps its own register budget
				if (V == 0 && a.vec == 4)
				{
					if constexpr (V == 0)
						gradient_tile<4, K, TY>(a, (int)t);
				}
				else if (a.vec == flat4_vec)
					gradient_tile_flat4<K, TY>(a, (int)t);
				else
					gradient_tile<1, K, TY>(a, (int)t);
			}
		}
	}
}

// Calibration kernel: float4 copy of n4 16-byte elements.  Its bandwidth is the practical ceiling the
// solver kernels are compared against, and its counter readings scale the traffic counters.
__global__ __launch_bounds__(256) void copy_kernel(float4* __restrict__ dst, const float4* __restrict__ src,
												   int64_t n4)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
		dst[i] = src[i];
}

void launch_copy(bfvm_ctx* ctx)
{
	const int64_t bytes = (int64_t)ctx->mesh.V * ctx->mesh.S * (ctx->storage == BFVM_STORAGE_F32 ? 4 : 2);
	const int64_t n4 = bytes / 16;
	if (n4 == 0)
		return;
	// one 16-byte element per thread in 256-thread workgroups, one pass over the array (up to 2^30 workgroups):
	// the launch shape that moved the most bytes per second on MI355X (6.1 TB/s against 5.2-5.8 TB/s for
	// persistent grid-stride launches and 5.8 TB/s for 1024-thread workgroups)
	const int64_t blocks = std::min<int64_t>(int64_t(1) << 30, (n4 + 255) / 256);
	hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (float4*)ctx->grad,
					   (const float4*)ctx->dens, n4);
	BFVM_HIP(hipGetLastError());
}

// Stacked tiles in 16-bit storage with 4 substrates per lane give S / 4 lanes at each end of a wave to halo
// columns, so they need S / 4 < 32; otherwise the side-by-side layout is used
static bool stacked_gradient_ok(const bfvm_ctx* ctx)
{
	return ctx->storage == BFVM_STORAGE_F32 || ctx->mesh.S % 4 != 0 || 2 * (ctx->mesh.S / 4) < wave_lanes;
}
static bool stacked_gradient_ok(const bfvm_batch* b)
{
	for (const bfvm_ctx* c : b->ctx)
		if (!stacked_gradient_ok(c))
			return false;
	return true;
}

// Tiling of one 3-D field with `rows` y rows per thread: z is split so that there are about `target` tiles
// (workgroups).  Stacked tiles: 64 columns by 4 * rows rows per workgroup; otherwise 256 columns by `rows`.
// flat: one-element instances take gradient_tile_flat4 (vec = flat4_vec, 4 row positions per thread)
static gradient_args make_gradient_args(const bfvm_ctx* ctx, int64_t target, int rows, bool stacked = false,
										bool flat = false)
{
	const auto& m = ctx->mesh;
	gradient_args a;
	a.grad = ctx->grad;
	a.dens = ctx->dens;
	a.inv = ctx->dinv;
	a.Y = m.grid[1];
	a.Z = m.grid[2];
	a.S = m.S;
	a.XS = m.grid[0] * m.S;
	a.vec = m.S % 4 == 0 ? 4 : (flat && !stacked ? flat4_vec : 1);
	// stacked 16-bit 4-substrate tiles: S / 4 halo lanes at each end of a wave
	const int halo = stacked && a.vec == 4 && ctx->storage != BFVM_STORAGE_F32 ? m.S / 4 : 0;
	if (2 * halo >= wave_lanes)
		throw std::logic_error("stacked gradient tiles: too many substrates for the halo lanes");
	const int width = stacked ? wave_lanes - 2 * halo : grad_tile;
	const int band = stacked ? 4 * rows : rows;
	const int per_lane = a.vec == flat4_vec ? 4 : a.vec;
	a.tiles_per_row = ((a.XS + per_lane - 1) / per_lane + width - 1) / width;
	const int64_t row_tiles = (int64_t)((a.Y + band - 1) / band) * a.tiles_per_row;
	int zchunks = (int)std::max<int64_t>(1, std::min<int64_t>(a.Z, (target + row_tiles - 1) / row_tiles));
	a.zchunk = (a.Z + zchunks - 1) / zchunks;
	a.zchunks = (a.Z + a.zchunk - 1) / a.zchunk;
	a.ntiles = (int)(row_tiles * a.zchunks);
	a.bx = 1.f / m.vs[0];
	a.cx = 1.f / (m.vs[0] * 2);
	a.by = 1.f / m.vs[1];
	a.cy = 1.f / (m.vs[1] * 2);
	a.bz = 1.f / m.vs[2];
	a.cz = 1.f / (m.vs[2] * 2);
	a.VS = (int64_t)m.V * m.S;
	return a;
}

void free_batch_gradient(bfvm_batch* b)
{
	auto& gp = b->grad;
	if (gp.inst)
		BFVM_HIP(hipFree(gp.inst));
	for (auto* items : gp.items)
		if (items)
			BFVM_HIP(hipFree(items));
	gp = bfvm_batch::gradient_plan {};
}

// Tiles of 16 z planes; items of b->tuning.gradient_tiles tiles (neighbouring rows of one z range, one after
// the other in one workgroup).
void prepare_batch_gradient(bfvm_batch* b)
{
	auto& gp = b->grad;
	std::vector<gradient_args> args(b->ctx.size());
	std::vector<batch_item> items[3]; // one element per thread, 4 substrates per thread, all
	for (size_t i = 0; i < b->ctx.size(); i++)
	{
		const bfvm_ctx* ctx = b->ctx[i];
		if (ctx->mesh.dims != 3)
		{
			gp.single.push_back((int)i);
			args[i] = gradient_args {};
			continue;
		}
		// tiles of `zc` z planes: ask for (row tiles) x ceil(Z / zc)
		const int rows = b->tuning.gradient_rows, zc = std::max(1, b->tuning.gradient_zchunk);
		const bool stacked = b->tuning.gradient_stacked != 0 && stacked_gradient_ok(b);
		const bool flat = true; // one-element instances as 4 row positions per thread
		const gradient_args probe = make_gradient_args(ctx, 1, rows, stacked, flat);
		const int64_t row_tiles = (int64_t)probe.ntiles;
		args[i] = make_gradient_args(ctx, row_tiles * ((ctx->mesh.grid[2] + zc - 1) / zc), rows, stacked, flat);
		const int m = args[i].vec == 4 ? 1 : 0;
		const int per_item = std::max(1, b->tuning.gradient_tiles);
		for (int64_t t0 = 0; t0 < args[i].ntiles; t0 += per_item)
		{
			const batch_item w { (int32_t)i, 0, t0, std::min<int64_t>(args[i].ntiles, t0 + per_item) };
			items[m].push_back(w);
			items[2].push_back(w);
		}
	}
	if (items[2].empty())
		return;
	BFVM_HIP(hipMalloc(&gp.inst, sizeof(gradient_args) * args.size()));
	BFVM_HIP(hipMemcpy(gp.inst, args.data(), sizeof(gradient_args) * args.size(), hipMemcpyHostToDevice));
	for (int m = 0; m < 3; m++)
	{
		if (items[m].empty())
			continue;
		BFVM_HIP(hipMalloc(&gp.items[m], sizeof(batch_item) * items[m].size()));
		BFVM_HIP(hipMemcpy(gp.items[m], items[m].data(), sizeof(batch_item) * items[m].size(), hipMemcpyHostToDevice));
		gp.nitems[m] = (int)items[m].size();
		gp.blocks[m] = std::min(gp.nitems[m], b->tuning.gradient_blocks);
	}
}

template <int TY, bool STACKED>
static void launch_batched_gradient_t(bfvm_batch* b)
{
	const auto& gp = b->grad;
	by_storage(b->storage, [&](auto kind) {
		constexpr int K = decltype(kind)::value;
		// FP32: one launch over both thread layouts unless tuning split it; 16-bit: one launch per layout
		if (K == store_f32 && !b->tuning.gradient_split)
		{
			if (gp.nitems[2])
				hipLaunchKernelGGL((gradient_batched_kernel<K, 0, TY, STACKED>), dim3(gp.blocks[2]), dim3(grad_tile), 0,
								   b->stream, (const gradient_args*)gp.inst, gp.items[2], gp.nitems[2]);
			return;
		}
		if (gp.nitems[1])
			hipLaunchKernelGGL((gradient_batched_kernel<K, 4, TY, STACKED>), dim3(gp.blocks[1]), dim3(grad_tile), 0,
							   b->stream,
							   (const gradient_args*)gp.inst, gp.items[1], gp.nitems[1]);
		if (gp.nitems[0])
			hipLaunchKernelGGL((gradient_batched_kernel<K, 1, TY, STACKED>), dim3(gp.blocks[0]), dim3(grad_tile), 0,
							   b->stream,
							   (const gradient_args*)gp.inst, gp.items[0], gp.nitems[0]);
	});
	BFVM_HIP(hipGetLastError());
}

void launch_batched_gradient(bfvm_batch* b)
{
	const bool st = b->tuning.gradient_stacked != 0 && stacked_gradient_ok(b);
	if (b->tuning.gradient_rows == 8)
		st ? launch_batched_gradient_t<8, true>(b) : launch_batched_gradient_t<8, false>(b);
	else
		st ? launch_batched_gradient_t<4, true>(b) : launch_batched_gradient_t<4, false>(b);
}

void launch_gradient(bfvm_ctx* ctx)
{
	const auto& m = ctx->mesh;
	const int64_t total = (int64_t)m.V * m.S;
	float b[3] = { 0, 0, 0 }, c[3] = { 0, 0, 0 };
	for (int d = 0; d < m.dims; d++)
	{
		b[d] = 1.f / m.vs[d];
		c[d] = 1.f / (m.vs[d] * 2);
	}
	if (m.dims == 3)
	{
		const int64_t target = ctx->gradient_groups > 0 ? ctx->gradient_groups : 16384;
		by_storage(ctx->storage, [&](auto k) {
			constexpr int K = decltype(k)::value;
			const bool v4 = m.S % 4 == 0;
			const bool st = ctx->gradient_stacked && stacked_gradient_ok(ctx);
			const gradient_args a =
				make_gradient_args(ctx, target, v4 ? gradient_rows<4, K>() : gradient_rows<1, K>(), st);
			const int per_xcd = (a.ntiles + num_xcds - 1) / num_xcds;
			auto kernel = v4 ? (st ? gradient3d_kernel<4, K, true> : gradient3d_kernel<4, K, false>)
							 : (st ? gradient3d_kernel<1, K, true> : gradient3d_kernel<1, K, false>);
			hipLaunchKernelGGL(kernel, dim3((unsigned)(per_xcd * num_xcds)), dim3(grad_tile), 0, ctx->stream, a,
							   per_xcd, ctx->gradient_order);
		});
		BFVM_HIP(hipGetLastError());
		return;
	}
	by_storage(ctx->storage, [&](auto k) {
		hipLaunchKernelGGL((gradient_kernel<decltype(k)::value>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
						   ctx->stream, ctx->grad, ctx->dens, ctx->dinv, m.grid[0], m.grid[1], m.grid[2], m.S, m.dims,
						   b[0], c[0], b[1], c[1], b[2], c[2], total);
	});
	BFVM_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------------
// FP32 <-> storage conversions (uploads and downloads of the densities)
// ------------------------------------------------------------------------------------------------

template <int K>
__global__ __launch_bounds__(256) void pack_kernel(void* __restrict__ dens, const float* __restrict__ src,
												   const float* __restrict__ scale, int S, int64_t offset, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const int s = (int)((offset + i) % S);
	dstore<K, 1>(dens, offset + i, src[i] * scale[s]);
}

template <int K>
__global__ __launch_bounds__(256) void unpack_kernel(float* __restrict__ dst, const void* __restrict__ dens,
													 const float* __restrict__ inv, int S, int64_t offset, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const int s = (int)((offset + i) % S);
	dst[i] = dload<K, 1>(dens, offset + i) * inv[s];
}

// per-substrate max |x| as the bit pattern of a non-negative float (which orders like an unsigned integer)
__global__ __launch_bounds__(256) void max_abs_kernel(const float* __restrict__ src, int S, int64_t n,
													  unsigned* __restrict__ out)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		const float a = fabsf(src[i]);
		if (a > 0.f && a <= 3.4e38f)
			atomicMax(&out[i % S], __float_as_uint(a));
	}
}

// FP16 storage: stored values of substrate s times factor[s] (a power of two, so exact while the result stays
// normal); grid-stride over the whole densities array
__global__ __launch_bounds__(256) void rescale_f16_kernel(_Float16* __restrict__ dens, const float* __restrict__ factor,
														  int S, int64_t n)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		dens[i] = (_Float16)((float)dens[i] * factor[i % S]);
}

void launch_rescale_f16(bfvm_ctx* ctx, const float* factor, hipStream_t stream)
{
	const int64_t n = (int64_t)ctx->mesh.V * ctx->mesh.S;
	const unsigned blocks = (unsigned)std::min<int64_t>(8192, (n + 255) / 256);
	hipLaunchKernelGGL(rescale_f16_kernel, dim3(blocks), dim3(256), 0, stream, (_Float16*)ctx->dens, factor,
					   ctx->mesh.S, n);
	BFVM_HIP(hipGetLastError());
}

__global__ void advance_rounding_kernel(unsigned* counter)
{
	if (threadIdx.x == 0)
		*counter = *counter + 1;
}

void launch_advance_rounding(bfvm_ctx* ctx)
{
	hipLaunchKernelGGL(advance_rounding_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->rounding_step);
	BFVM_HIP(hipGetLastError());
}

void launch_pack(bfvm_ctx* ctx, const float* src, int64_t offset, int64_t n)
{
	by_storage(ctx->storage, [&](auto k) {
		hipLaunchKernelGGL((pack_kernel<decltype(k)::value>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
						   ctx->stream, ctx->dens, src, ctx->dscale, ctx->mesh.S, offset, n);
	});
	BFVM_HIP(hipGetLastError());
}

void launch_unpack(bfvm_ctx* ctx, float* dst, int64_t offset, int64_t n)
{
	by_storage(ctx->storage, [&](auto k) {
		hipLaunchKernelGGL((unpack_kernel<decltype(k)::value>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
						   ctx->stream, dst, ctx->dens, ctx->dinv, ctx->mesh.S, offset, n);
	});
	BFVM_HIP(hipGetLastError());
}

void launch_max_abs(bfvm_ctx* ctx, const float* src, int64_t n, unsigned* out)
{
	hipLaunchKernelGGL(max_abs_kernel, dim3(4096), dim3(256), 0, ctx->stream, src, ctx->mesh.S, n, out);
	BFVM_HIP(hipGetLastError());
}

} // namespace bfvm
