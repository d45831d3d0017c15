// LOD implicit diffusion on gfx950: Thomas coefficient setup and the x/y/z sweeps.
//
// The system of one line (host/diffusion_solver.h:359-390, coefficients host/diffusion_solver.cpp:72-142):
//   a x_{i-1} + b_i x_i + c x_{i+1} = d_i,  a = c = -dt*D/dx^2,
//   b_end = 1 + decay*dt/dims + dt*D/dx^2,  b_mid = 1 + decay*dt/dims + 2*dt*D/dx^2,
// and for a line of one voxel b = 1 + decay*dt/dims (host/diffusion_solver.cpp:76-85).
//
// Two sweep algorithms:
//  * exact: one thread per line, the reference's sequential recurrence with its FP32 coefficients in its
//    operation order (host/diffusion_solver.cpp:297-334, 378-480); bit-identical to the host path.
//  * partitioned: the line is split into P blocks held in registers, one thread per (line, block).  Each
//    block is reduced to its two end unknowns by the modified Thomas algorithm (the PaScaL_TDMA scheme the
//    reference precomputes in host/diffusion_solver.cpp:144-295 for its 2-D device kernels), the 2P-unknown
//    reduced system is solved per line from LDS, and the block interiors are reconstructed.  Every element
//    is read once and written once.  Coefficients are derived in FP64 on the GPU and rounded to FP32.
//
// Both apply the Dirichlet boundary faces in their store epilogue (the Dirichlet pass that follows each
// sweep in host/diffusion_solver.cpp:529,538,543); the interior list is applied by a separate kernel.
#include <algorithm>
#include <stdexcept>

#include "bfvm_internal.hpp"

namespace bfvm {

// ------------------------------------------------------------------------------------------------
// coefficients
// ------------------------------------------------------------------------------------------------

// Reference-order FP32 coefficients of one axis: host/diffusion_solver.cpp:72-142 (copies = 1).
// One thread per substrate; the recurrence over i is sequential.
__global__ void exact_coefficients_kernel(float* __restrict__ b, float* __restrict__ c, float* __restrict__ e,
										  const float* __restrict__ D, const float* __restrict__ decay, float dt,
										  int shape, int dims, int n, int S)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= S)
		return;
	if (n == 1)
	{
		b[s] = 1 / (1 + decay[s] * dt / dims);
		return;
	}
	const float cs = -dt * D[s] / (shape * shape);
	c[s] = cs;
	const float b_end = 1 + decay[s] * dt / dims + dt * D[s] / (shape * shape);
	const float b_mid = 1 + decay[s] * dt / dims + 2 * dt * D[s] / (shape * shape);
	float prev = 1 / b_end;
	b[s] = prev;
	for (int i = 1; i < n; i++)
	{
		const float bi = (i == n - 1) ? b_end : b_mid;
		const float cur = 1 / (bi - cs * cs * prev);
		b[(size_t)i * S + s] = cur;
		e[(size_t)(i - 1) * S + s] = cs * prev;
		prev = cur;
	}
}

// Partitioned (modified Thomas) coefficients of one axis in FP64, rounded to FP32.  One thread per substrate.
// Block of m elements, local j, diagonal B_j (B_end at the line's ends, B_mid elsewhere), off-diagonal A:
//   forward  j<2: r=1/B_j, alpha_j=A r, gamma_j=A r;  j>=2: r=1/(B_j - A gamma_{j-1}),
//            alpha_j=-A alpha_{j-1} r, gamma_j=A r       -> row j: alpha_j x_0 + x_j + gamma_j x_{j+1} = delta_j
//   backward j=m-3..1: alpha'_j = alpha_j - gamma_j alpha'_{j+1}, gamma'_j = -gamma_j gamma'_{j+1}
//            j=0 (m>=3): rho = 1/(1 - gamma_0 alpha'_1), alpha'_0 = rho alpha_0, gamma'_0 = -rho gamma_0 gamma'_1
//                                                         -> row j: alpha'_j x_0 + x_j + gamma'_j x_{m-1} = delta'_j
// The reduced system of the line has unknowns u_{2k} = first, u_{2k+1} = last of block k, with rows
//   u_{2k}:   alpha'_0 u_{2k-1} + u_{2k} + gamma'_0 u_{2k+1}
//   u_{2k+1}: alpha_{m-1} u_{2k} + u_{2k+1} + gamma_{m-1} u_{2k+2}
// i.e. unit diagonal, sub-diagonal L_q and super-diagonal U_q.  It is solved from both ends at once (twisted
// factorization): rows q <= h = P-1 are eliminated downwards, rows q > h upwards, and the two halves meet in
// the 2x2 system of rows h, h+1:
//   top    q <= h: inv_q  = 1 / (1 - L_q c'_{q-1}),  c'_q  = U_q inv_q;   stored f0 = L_q inv_q,  f1 = inv_q,  f2 = c'_q
//   bottom q >  h: jinv_q = 1 / (1 - U_q a''_{q+1}), a''_q = L_q jinv_q;  stored f0 = U_q jinv_q, f1 = jinv_q, f2 = a''_q
//   meeting factor mu = 1 / (1 - c'_h a''_{h+1})
// Table layout: coef[f][type][j][s], f = 0 r, 1 gamma, 2 alpha' (rho in the j=0 slot), 3 gamma';
// red: 16-byte records {f0, f1, f2, 0} at [q][s], then mu[s].
constexpr int max_types = 4;

// floats of the reduced-system table of a line of P blocks
__host__ __device__ constexpr int reduced_table_floats(int P, int S) { return 4 * 2 * P * S + S; }
// LDS of a partitioned sweep workgroup, in floats: the coefficient table [4][ntypes][mb][Sp], the reduced-system
// table, A and the storage scale [Sp] each, then (from the returned offset, 16-byte aligned) the reduced rows.  Sp is
// S, or S + V - 1 when substrate vectors wrap (rows extended periodically so V consecutive entries from any
// substrate are the wrapped substrates).
__host__ __device__ constexpr int part_rows_offset(int ntypes, int mb, int P, int S, int Sp)
{
	return (4 * ntypes * mb * Sp + reduced_table_floats(P, S) + 2 * Sp + 3) / 4 * 4;
}

__global__ void partition_coefficients_kernel(float* __restrict__ coef, float* __restrict__ red,
											  float* __restrict__ A_out, const float* __restrict__ D,
											  const float* __restrict__ decay, float dt, int shape, int dims, int n,
											  int S, int P, int base, int rem, int ntypes, int mb)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= S)
		return;
	const double ddt = dt, dD = D[s], dl = decay[s];
	const double h2 = (double)shape * (double)shape;
	const double A = -ddt * dD / h2;
	const double B_end = 1.0 + dl * ddt / dims + ddt * dD / h2;
	const double B_mid = 1.0 + dl * ddt / dims + 2.0 * ddt * dD / h2;
	A_out[s] = (float)A;

	double ap0[max_types], gp0[max_types], al_last[max_types], ga_last[max_types];
	const size_t fstride = (size_t)ntypes * mb * S;

	for (int t = 0; t < ntypes; t++)
	{
		int m;
		bool first, last;
		if (P == 1)
		{
			m = n;
			first = last = true;
		}
		else
		{
			first = t == 0;
			last = t == 3;
			m = (t == 0) ? base + (rem > 0 ? 1 : 0) : (t == 1 ? base + 1 : base);
		}
		// forward: alpha, gamma kept in doubles for the backward pass
		double alpha[129], gamma[129];
		for (int j = 0; j < m; j++)
		{
			const bool end = (first && j == 0) || (last && j == m - 1);
			const double Bj = end ? B_end : B_mid;
			double r;
			if (j < 2)
			{
				r = 1.0 / Bj;
				alpha[j] = A * r;
			}
			else
			{
				r = 1.0 / (Bj - A * gamma[j - 1]);
				alpha[j] = -A * alpha[j - 1] * r;
			}
			gamma[j] = A * r;
			coef[0 * fstride + ((size_t)t * mb + j) * S + s] = (float)r;
			coef[1 * fstride + ((size_t)t * mb + j) * S + s] = (float)gamma[j];
		}
		// backward
		double apn = 0, gpn = 0; // alpha', gamma' of row j+1
		for (int j = m - 2; j >= 1; j--)
		{
			double apj, gpj;
			if (j == m - 2)
			{
				apj = alpha[j];
				gpj = gamma[j];
			}
			else
			{
				apj = alpha[j] - gamma[j] * apn;
				gpj = -gamma[j] * gpn;
			}
			coef[2 * fstride + ((size_t)t * mb + j) * S + s] = (float)apj;
			coef[3 * fstride + ((size_t)t * mb + j) * S + s] = (float)gpj;
			apn = apj;
			gpn = gpj;
		}
		if (m >= 3)
		{
			const double rho = 1.0 / (1.0 - gamma[0] * apn);
			ap0[t] = rho * alpha[0];
			gp0[t] = -rho * gamma[0] * gpn;
			coef[2 * fstride + ((size_t)t * mb + 0) * S + s] = (float)rho;
		}
		else
		{
			ap0[t] = alpha[0];
			gp0[t] = gamma[0];
			coef[2 * fstride + ((size_t)t * mb + 0) * S + s] = 1.0f;
		}
		coef[3 * fstride + ((size_t)t * mb + 0) * S + s] = 0.0f;
		al_last[t] = alpha[m - 1];
		ga_last[t] = gamma[m - 1];
	}

	// reduced system, twisted factorization with unit diagonal
	const int N = 2 * P, h = P - 1;
	auto record = [&](int q, double f0, double f1, double f2) {
		float* rec = red + ((size_t)q * S + s) * 4;
		rec[0] = (float)f0;
		rec[1] = (float)f1;
		rec[2] = (float)f2;
		rec[3] = 0.f;
	};
	auto off_diagonals = [&](int q, double& L, double& U) {
		const int k = q / 2;
		const int t = (P == 1) ? 0 : (k == 0 ? 0 : (k == P - 1 ? 3 : (k < rem ? 1 : 2)));
		if ((q & 1) == 0)
		{
			L = (k > 0) ? ap0[t] : 0.0;
			U = gp0[t];
		}
		else
		{
			L = al_last[t];
			U = (k < P - 1) ? ga_last[t] : 0.0;
		}
	};
	double cprev = 0; // c'_{q-1}
	for (int q = 0; q <= h; q++)
	{
		double L, U;
		off_diagonals(q, L, U);
		const double inv = q == 0 ? 1.0 : 1.0 / (1.0 - L * cprev);
		const double cp = U * inv;
		record(q, L * inv, inv, cp);
		cprev = cp;
	}
	double anext = 0; // a''_{q+1}
	for (int q = N - 1; q > h; q--)
	{
		double L, U;
		off_diagonals(q, L, U);
		const double jinv = q == N - 1 ? 1.0 : 1.0 / (1.0 - U * anext);
		const double a2 = L * jinv;
		record(q, U * jinv, jinv, a2);
		anext = a2;
	}
	red[(size_t)N * S * 4 + s] = (float)(1.0 / (1.0 - cprev * anext));
}

// ------------------------------------------------------------------------------------------------
// Dirichlet faces in a sweep's store epilogue
// ------------------------------------------------------------------------------------------------

// Per-thread view of the six faces for one substrate and a line along `axis` with the other two
// coordinates fixed: the override from faces of axes before `axis` (lo), after it (hi), and the two faces
// of `axis` itself.  Face precedence is x-, x+, y-, y+, z-, z+ (host/dirichlet_solver.cpp:91-110).
struct face_epilogue
{
	bool lo_set = false, hi_set = false, cmin = false, cmax = false;
	float lo_val = 0, hi_val = 0, vmin = 0, vmax = 0;

	// `scale`: the storage scale of substrate s, for sweeps that solve on stored (scaled) values
	__device__ void init(const float* __restrict__ fval, const uint8_t* __restrict__ fcond, int S, int s, int axis,
						 const int* coord, const int* grid, int dims, float scale = 1.f)
	{
		if (!fval)
			return;
		for (int d = 0; d < dims; d++)
		{
			for (int mx = 0; mx < 2; mx++)
			{
				const int f = 2 * d + mx;
				const bool cond = fcond[f * S + s] != 0;
				const float val = fval[f * S + s] * scale;
				if (d == axis)
				{
					if (mx == 0)
					{
						cmin = cond;
						vmin = val;
					}
					else
					{
						cmax = cond;
						vmax = val;
					}
					continue;
				}
				const bool on = mx == 0 ? coord[d] == 0 : coord[d] == grid[d] - 1;
				if (on && cond)
				{
					if (d < axis)
					{
						lo_set = true;
						lo_val = val;
					}
					else
					{
						hi_set = true;
						hi_val = val;
					}
				}
			}
		}
	}

	__device__ __forceinline__ float apply(float v, int i, int n) const
	{
		if (lo_set)
			v = lo_val;
		if (i == 0 && cmin)
			v = vmin;
		if (i == n - 1 && cmax)
			v = vmax;
		if (hi_set)
			v = hi_val;
		return v;
	}
};

struct sweep_geometry
{
	void* data;		   // densities in the storage form
	const float* scale; // [S] storage scale (16-bit forms)
	const float* inv;	// [S] 1 / scale
	const unsigned* rounding; // step counter keying the stochastic rounding (16-bit forms)
	int X, Y, Z, S, dims;
	const float* fval;
	const uint8_t* fcond;
};

// Line geometry per axis: the element (line, i) lives at base + i*stride.
//  x: line = (row r = z*Y + y, s), stride S
//  y: line = (z, q) with q = x*S + s, stride X*S
//  z: line = q in [0, Y*X*S), stride Y*X*S
template <int AXIS>
__device__ __forceinline__ void line_coords(const sweep_geometry& g, int64_t line, int64_t& base, int& s, int* coord)
{
	const int64_t XS = (int64_t)g.X * g.S;
	if constexpr (AXIS == 0)
	{
		const int64_t row = line / g.S;
		s = (int)(line % g.S);
		base = row * XS + s;
		coord[0] = 0;
		coord[1] = (int)(row % g.Y);
		coord[2] = (int)(row / g.Y);
	}
	else if constexpr (AXIS == 1)
	{
		const int64_t z = line / XS;
		const int64_t q = line % XS;
		s = (int)(q % g.S);
		base = z * (int64_t)g.Y * XS + q;
		coord[0] = (int)(q / g.S);
		coord[1] = 0;
		coord[2] = (int)z;
	}
	else
	{
		s = (int)(line % g.S);
		base = line;
		coord[0] = (int)((line / g.S) % g.X);
		coord[1] = (int)(line / XS);
		coord[2] = 0;
	}
}

template <int AXIS>
__device__ __forceinline__ int64_t line_stride(const sweep_geometry& g)
{
	if constexpr (AXIS == 0)
		return g.S;
	else if constexpr (AXIS == 1)
		return (int64_t)g.X * g.S;
	else
		return (int64_t)g.Y * g.X * g.S;
}

// ------------------------------------------------------------------------------------------------
// exact sweep: one thread per line (host/diffusion_solver.cpp:297-334 along x, :378-480 along y/z)
// ------------------------------------------------------------------------------------------------

template <int AXIS>
__global__ __launch_bounds__(256) void sweep_exact_kernel(sweep_geometry g, int64_t lines, int n,
														  const float* __restrict__ b, const float* __restrict__ c,
														  const float* __restrict__ e)
{
	const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (line >= lines)
		return;
	int64_t base;
	int s;
	int coord[3];
	line_coords<AXIS>(g, line, base, s, coord);
	const int64_t stride = line_stride<AXIS>(g);
	float* p = (float*)g.data + base; // FP32 storage only: the forward pass keeps its intermediates in memory
	const int grid[3] = { g.X, g.Y, g.Z };
	face_epilogue fe;
	fe.init(g.fval, g.fcond, g.S, s, AXIS, coord, grid, g.dims);

	float prev = p[0];
	for (int i = 1; i < n; i++)
	{
		float cur = p[i * stride];
		cur = cur - e[(size_t)(i - 1) * g.S + s] * prev;
		p[i * stride] = cur;
		prev = cur;
	}
	const float cs = c[s];
	float next = prev * b[(size_t)(n - 1) * g.S + s];
	p[(int64_t)(n - 1) * stride] = fe.apply(next, n - 1, n);
	for (int i = n - 2; i >= 0; i--)
	{
		float cur = p[i * stride];
		cur = (cur - cs * next) * b[(size_t)i * g.S + s];
		p[i * stride] = fe.apply(cur, i, n);
		next = cur;
	}
}

// Axis of length one: the sweep is x *= b (b = 1/(1 + decay*dt/dims)), then the faces.
template <int K>
__global__ __launch_bounds__(256) void sweep_single_kernel(sweep_geometry g, int axis, const float* __restrict__ b)
{
	const int64_t total = (int64_t)g.X * g.Y * g.Z * g.S;
	const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= total)
		return;
	const int s = (int)(idx % g.S);
	const int64_t v = idx / g.S;
	int coord[3] = { (int)(v % g.X), (int)((v / g.X) % g.Y), (int)(v / ((int64_t)g.X * g.Y)) };
	const int grid[3] = { g.X, g.Y, g.Z };
	face_epilogue fe;
	fe.init(g.fval, g.fcond, g.S, s, axis, coord, grid, g.dims);
	if constexpr (K == store_f32)
		((float*)g.data)[idx] = fe.apply(((float*)g.data)[idx] * b[s], 0, 1);
	else
	{
		const float d = dload<K, 1>(g.data, idx) * g.inv[s];
		dstore_sr<K, 1>(g.data, idx, fe.apply(d * b[s], 0, 1) * g.scale[s], rounding_key(g.rounding, 1 + axis));
	}
}

// ------------------------------------------------------------------------------------------------
// partitioned sweep
// ------------------------------------------------------------------------------------------------

// V consecutive substrates of one voxel row are handled by one thread as a vector (V = 4 when the
// substrate count is a multiple of 4): every global access is then a 16-byte load or store and every
// coefficient read a 16-byte LDS read (vec_of / vload / vstore in bfvm_internal.hpp).

struct part_args
{
	sweep_geometry g;
	int n, P, base, rem, ntypes, mb;
	const float* coef;
	const float* red;
	const float* A;
	int64_t lines;		 // number of lines (groups of V substrates)
	int lines_per_group; // lines handled by one workgroup iteration
	int64_t groups;
	int vec; // substrates per thread (V)
	// the two divisors of this axis's line decomposition (x: S/V and Y, y: X*S/V and S, z: S and X) as
	// multiply-shift pairs: n / d = (n * mul) >> shift for every n < 2^31 (the host checks lines * V < 2^31)
	uint32_t div_mul[2];
	int div_shift[2];
};

// n / d for n < 2^31 with (mul, shift) from fast_divisor(d): Granlund and Montgomery's multiply-shift division,
// mul = ceil(2^(31+l) / d) with 2^l >= d, which fits 32 bits and is exact for 31-bit numerators
__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t mul, int shift)
{
	return (uint32_t)(((uint64_t)n * mul) >> shift);
}

// Line geometry with V substrates per line.  x: line = (row, substrate group); y: line = (z, q0/V) with
// q0 = x*S + s0; z: line = q0/V over the whole (y, x, s) plane.  Line indices and plane positions stay below
// 2^31 (checked on the host), so the divisions are 32-bit multiply-shifts by precomputed divisors.
template <int AXIS, int V>
__device__ __forceinline__ void vline_coords(const part_args& a, int64_t line, int64_t& base, int& s0, int* coord)
{
	const sweep_geometry& g = a.g;
	const int64_t XS = (int64_t)g.X * g.S;
	const uint32_t l = (uint32_t)line;
	if constexpr (AXIS == 0)
	{
		const uint32_t SG = (uint32_t)(g.S / V);
		const uint32_t row = fdiv(l, a.div_mul[0], a.div_shift[0]);
		s0 = (int)(l - row * SG) * V;
		base = (int64_t)row * XS + s0;
		const uint32_t z = fdiv(row, a.div_mul[1], a.div_shift[1]);
		coord[0] = 0;
		coord[1] = (int)(row - z * (uint32_t)g.Y);
		coord[2] = (int)z;
	}
	else if constexpr (AXIS == 1)
	{
		const uint32_t XSV = (uint32_t)(XS / V);
		const uint32_t z = fdiv(l, a.div_mul[0], a.div_shift[0]);
		const uint32_t q0 = (l - z * XSV) * V;
		const uint32_t x = fdiv(q0, a.div_mul[1], a.div_shift[1]);
		s0 = (int)(q0 - x * (uint32_t)g.S);
		base = (int64_t)z * g.Y * XS + q0;
		coord[0] = (int)x;
		coord[1] = 0;
		coord[2] = (int)z;
	}
	else
	{
		const uint32_t q0 = l * V;
		const uint32_t xy = fdiv(q0, a.div_mul[0], a.div_shift[0]);
		s0 = (int)(q0 - xy * (uint32_t)g.S);
		base = q0;
		const uint32_t y = fdiv(xy, a.div_mul[1], a.div_shift[1]);
		coord[0] = (int)(xy - y * (uint32_t)g.X);
		coord[1] = (int)y;
		coord[2] = 0;
	}
}

// Thread t of a workgroup handles block k of line `line_local`:
//  x: t = (line_local / SG) * (SG*P) + k*SG + sg -> the lanes of one x row read SG*V consecutive floats per
//     block (SG = S/V substrate groups)
//  y, z: t = k * lines_per_group + line_local   -> lanes read consecutive 16-byte vectors of a plane row
template <int AXIS>
__device__ __forceinline__ void thread_slot(int t, int SG, int P, int lpg, int& line_local, int& k)
{
	if constexpr (AXIS == 0)
	{
		const int sg = t % SG;
		k = (t / SG) % P;
		line_local = (t / (SG * P)) * SG + sg;
		if (t >= lpg * P)
			k = -1;
	}
	else
	{
		line_local = t % lpg;
		k = t / lpg;
		if (k >= P)
			k = -1;
	}
}

// Substrate s and voxel coordinates cc of component c of a thread's substrate vector (first substrate s0,
// line coordinates `coord`).  Without wrapping the V substrates belong to one voxel.  With wrapping (y/z
// sweeps whose substrate count V does not divide) the vector holds V consecutive flattened plane positions,
// so component c is substrate (s0 + c) mod S of a voxel (s0 + c) / S further along x (and, along z, possibly
// in the next y row).
template <int AXIS, bool WRAP>
__device__ __forceinline__ void comp_coords(const sweep_geometry& g, const int* coord, int s0, int c, int& s, int* cc)
{
	cc[0] = coord[0];
	cc[1] = coord[1];
	cc[2] = coord[2];
	if constexpr (!WRAP)
		s = s0 + c;
	else
	{
		static_assert(AXIS != 0, "x lines hold whole voxels");
		const int t = s0 + c;
		s = t % g.S;
		const int adv = t / g.S;
		if constexpr (AXIS == 1)
			cc[0] += adv;
		else
		{
			const int xx = coord[0] + adv;
			cc[0] = xx % g.X;
			cc[1] += xx / g.X;
		}
	}
}

// Dirichlet faces for the elements a thread stores: the common case (an interior line away from its ends)
// stores the solved value; lines on a face of another axis, and the two ends of the line, resolve every
// face in the order x-, x+, y-, y+, z-, z+ (host/dirichlet_solver.cpp:91-110).
template <int AXIS, int V, bool WRAP, int K>
__device__ __forceinline__ typename vec_of<V>::type faces_slow(typename vec_of<V>::type x, const sweep_geometry& g,
															  int s0, const int* coord, int i,
															  const typename vec_of<V>::type& scale)
{
	const int grid[3] = { g.X, g.Y, g.Z };
	for (int c = 0; c < V; c++)
	{
		int s = s0 + c, cc[3];
		if constexpr (WRAP)
			comp_coords<AXIS, WRAP>(g, coord, s0, c, s, cc);
		float val = comp(x, c);
		bool face = false;
		for (int d = 0; d < g.dims; d++)
		{
			const int cd = d == AXIS ? i : (WRAP ? cc[d] : coord[d]);
			if (cd == 0 && g.fcond[(2 * d) * g.S + s])
			{
				val = g.fval[(2 * d) * g.S + s];
				face = true;
			}
			if (cd == grid[d] - 1 && g.fcond[(2 * d + 1) * g.S + s])
			{
				val = g.fval[(2 * d + 1) * g.S + s];
				face = true;
			}
		}
		// 16-bit forms: a face value takes the storage scale (solved values already carry it)
		if constexpr (K != store_f32)
			if (face)
				val *= comp(scale, c);
		set_comp(x, c, val);
	}
	return x;
}

// Workgroups of the 4-substrate kernels are capped at 512 threads so each thread may hold up to 256 VGPRs
// (a block of 16 float4 values plus coefficients does not fit the 128 of a 1024-thread workgroup).
constexpr int max_group_threads(int V) { return V == 4 ? 512 : 1024; }

// Twisted Thomas solve of one line's 2P-row reduced system, run by two adjacent lanes of a wave: half 0
// eliminates rows 1..h downwards, half 1 rows 2P-2..h+1 upwards (h = P-1), the lanes swap their last values,
// both resolve rows h and h+1, and each substitutes back towards its own end.  Rows r[q * rs] in LDS, in
// place; table records {f0, f1, f2} at tab[q * S] and mu (both offset to the line's substrate).  Each step is
// one multiply-add on the critical path, and the LDS reads of step i+1 are issued before the arithmetic of
// step i.  The chain is P steps each way instead of 2P for a one-ended solve.
__device__ __forceinline__ void twisted_solve(const float4* tab, float mu, int S, float* r, int rs, int P, int half)
{
#pragma clang fp contract(fast)
	const int h = P - 1;
	const int dr = half ? -rs : rs;
	const int dt = half ? -S : S;
	float* rp = r + (half ? (2 * P - 1) * rs : 0);
	const float4* tp = tab + (half ? (2 * P - 1) * S : 0);
	float y = *rp; // f1 is 1 in both end rows
	float Rn = 0.f;
	float2 Fn = make_float2(0.f, 0.f);
	if (h > 0)
	{
		Rn = rp[dr];
		Fn = *(const float2*)(tp + dt);
	}
#pragma unroll 1
	for (int i = 1; i <= h; i++)
	{
		rp += dr;
		tp += dt;
		const float Rq = Rn;
		const float2 Fq = Fn;
		if (i < h)
		{
			Rn = rp[dr];
			Fn = *(const float2*)(tp + dt);
		}
		y = Rq * Fq.y - Fq.x * y;
		*rp = y;
	}
	// rp, tp are at row h (top lane) or h+1 (bottom lane), y is that row's eliminated value; both lanes solve
	//   u_h = (top - c'_h bottom) mu,  u_{h+1} = bottom - a''_{h+1} u_h
	const float other = __shfl_xor(y, 1);
	const float top = half ? other : y;
	const float bottom = half ? y : other;
	const float uh = (top - tab[h * S].z * bottom) * mu;
	y = half ? bottom - tp->z * uh : uh;
	*rp = y;
	// back substitution towards the lane's end: u_q = r_q - f2_q u_{q -+ 1}
	float Cn = 0.f;
	if (h > 0)
	{
		Rn = rp[-dr];
		Cn = tp[-dt].z;
	}
#pragma unroll 1
	for (int i = 1; i <= h; i++)
	{
		rp -= dr;
		tp -= dt;
		const float Rq = Rn, Cq = Cn;
		if (i < h)
		{
			Rn = rp[-dr];
			Cn = tp[-dt].z;
		}
		y = Rq - Cq * y;
		*rp = y;
	}
}

// Groups g_first, g_first + g_step, ... < g_end of one partitioned sweep, by one workgroup.  The caller
// synchronises the workgroup before reusing `lds`.  FULL: every block holds M - 1 or M elements
// (a.base >= M - 1), so only the last element's presence differs between threads and the other per-element
// guards are known at compile time.
template <int M, int AXIS, int V, int K, bool PF = (K != store_f32 && AXIS != 0), bool NT = (K == store_f32 && AXIS != 0),
		  bool FULL = false>
__device__ __forceinline__ void sweep_groups(const part_args& a, float* lds, int64_t g_first, int64_t g_end,
											 int64_t g_step)
{
	// this path is held to a tolerance, not to the reference's operation order: multiply-adds may fuse
#pragma clang fp contract(fast)
	using vec = typename vec_of<V>::type;
	const sweep_geometry& g = a.g;
	int S = g.S; // redefined opaquely in every group iteration, below
	const int P = a.P;
	const int lpg = a.lines_per_group;
	const int ncoef = 4 * a.ntypes * a.mb * S;
	const int nred = reduced_table_floats(P, S);
	float* lcoef = lds;
	float* lred = lds + ncoef;
	// per-substrate A and storage scale, so the group loop reads no global memory besides the densities
	float* lA = lred + nred;
	float* lscale = lA + S;
	vec* rv = (vec*)(lds + part_rows_offset(a.ntypes, a.mb, P, S, S)); // [2P][lpg]

	for (int i = threadIdx.x; i < ncoef; i += blockDim.x)
		lcoef[i] = a.coef[i];
	for (int i = threadIdx.x; i < nred; i += blockDim.x)
		lred[i] = a.red[i];
	for (int i = threadIdx.x; i < S; i += blockDim.x)
	{
		lA[i] = a.A[i];
		lscale[i] = K == store_f32 ? 1.f : g.scale[i];
	}

	int line_local, k;
	thread_slot<AXIS>(threadIdx.x, S / V, P, lpg, line_local, k);
	int64_t stride = line_stride<AXIS>(g); // redefined opaquely in every group iteration, below
	const int grid[3] = { g.X, g.Y, g.Z };
	const size_t fstride = (size_t)a.ntypes * a.mb * S;
	const uint32_t key = K == store_f32 ? 0u : rounding_key(g.rounding, 1 + AXIS);

	int start = 0, mk = 0, type = 0;
	if (k >= 0)
	{
		start = k * a.base + min(k, a.rem);
		mk = a.base + (k < a.rem ? 1 : 0);
		type = (P == 1) ? 0 : (k == 0 ? 0 : (k == P - 1 ? 3 : (k < a.rem ? 1 : 2)));
	}
	// element j of the thread's block exists / is the block's last
	auto has = [&](int j) { return (FULL && j < M - 1) || j < mk; };
	auto last = [&](int j) { return (!FULL || j >= M - 2) && j == mk - 1; };
	__syncthreads();

	// PF (16-bit forms): load the next group's block while this one is solved.  A latency-bound sweep keeps a
	// fixed number of loads in flight, and 8-byte lanes bring half the bytes of FP32 ones per load.  Measured on
	// 400^3 x 8 BF16: y 1.54 -> 1.02 ms, z 1.94 -> 1.11 ms; x 1.04 -> 1.16 ms (there the lost wave of occupancy
	// costs more), while the batch's 4-wide x sweeps gain.
	constexpr bool prefetch = PF;
	typename svec_of<K, V>::type pre[prefetch ? M : 1];
	auto fetch = [&](int64_t grp) {
		const int64_t ln = grp * lpg + line_local;
		if (k >= 0 && ln < a.lines)
		{
			int64_t b = 0;
			int sx = 0;
			int cd[3];
			vline_coords<AXIS, V>(a, ln, b, sx, cd);
			const int64_t o = b + (int64_t)start * stride;
#pragma unroll
			for (int j = 0; j < M; j++)
				if (has(j))
					pre[j] = draw<K, V>(g.data, o + j * stride);
		}
	};
	if constexpr (prefetch)
		if (g_first < g_end)
			fetch(g_first);

	for (int64_t group = g_first; group < g_end; group += g_step)
	{
		// The unrolled element loops address memory at j * stride and the LDS tables at j * S for every j; held
		// across the group loop those 16 + 16 uniform offsets overflow the scalar registers and spill into
		// vector-register lanes (hundreds of SGPR spills per kernel).  Marking stride and S as redefined here
		// makes them recomputed per group instead: measured on 640^3 x 16 BF16 (autotuned) sweeps 8.7 / 9.3 /
		// 10.0 -> 7.5 / 8.7 / 8.5 ms, and 10-12 fewer VGPRs in every 4-substrate kernel.
		asm volatile("" : "+s"(stride), "+s"(S));
		const int64_t line = group * lpg + line_local;
		const bool active = k >= 0 && line < a.lines;
		int64_t base = 0;
		int s0 = 0;
		int coord[3] = { 0, 0, 0 };
		if (active)
			vline_coords<AXIS, V>(a, line, base, s0, coord);
		const int64_t off0 = base + (int64_t)start * stride; // element offset of the block's first value
		const float* cr = lcoef + 0 * fstride + (size_t)type * a.mb * S + s0;
		const float* cg = lcoef + 1 * fstride + (size_t)type * a.mb * S + s0;
		const float* ca = lcoef + 2 * fstride + (size_t)type * a.mb * S + s0;
		const float* cgp = lcoef + 3 * fstride + (size_t)type * a.mb * S + s0;

		vec v[M];
		if constexpr (prefetch)
		{
			if (active)
			{
#pragma unroll
				for (int j = 0; j < M; j++)
					if (has(j))
						v[j] = dcvt<K, V>(pre[j]);
			}
			if (group + g_step < g_end)
				fetch(group + g_step);
		}
		if (active)
		{
			if constexpr (!prefetch)
			{
#pragma unroll
				for (int j = 0; j < M; j++)
					if (has(j))
					{
						// NT: the FP32 y/z inputs are read once and streamed past the caches like the outputs
						// (400^3 x 8: y 0.881 -> 0.856 ms, z 0.941 -> 0.925 ms)
						if constexpr (NT)
							v[j] = __builtin_nontemporal_load((const vec*)((const float*)g.data + off0 + j * stride));
						else
							v[j] = dload<K, V>(g.data, off0 + j * stride);
					}
			}
			// 16-bit forms solve on the stored values: the scale is a power of two, so solving on scaled values and
			// scaling the solution give the same bits, and only the Dirichlet values set below need the scale
			// forward elimination inside the block: v_j = r_j (v_j - A v_{j-1}), written as r_j v_j - (r_j A) v_{j-1}
			// so that the recurrence carries one multiply-add
			const vec As = vload<V>(lA + s0);
#pragma unroll
			for (int j = 0; j < M; j++)
			{
				if (has(j))
				{
					const vec r = vload<V>(cr + j * S);
					if (j < 2)
						v[j] = r * v[j];
					else
						v[j] = r * v[j] - (r * As) * v[j - 1];
				}
			}
			// backward elimination towards the block's last unknown
#pragma unroll
			for (int j = M - 3; j >= 1; j--)
				if (has(j + 2))
					v[j] = v[j] - vload<V>(cg + j * S) * v[j + 1];
			if (has(2))
				v[0] = vload<V>(ca) * (v[0] - vload<V>(cg) * v[1]);
			rv[(2 * k) * lpg + line_local] = v[0];
#pragma unroll
			for (int j = 1; j < M; j++)
				if (last(j))
					rv[(2 * k + 1) * lpg + line_local] = v[j];
		}
		__syncthreads();
		// reduced systems: one pair of adjacent threads per (line, substrate), solving from both ends.  On every
		// axis the first substrate of line l is (l V) mod S (the lines of a row, column or plane position run
		// over the substrate groups), and the group's share of it is uniform across the workgroup.
		const int group_s = (int)((group * lpg * V) % S);
		for (int pr = threadIdx.x >> 1; threadIdx.x < (blockDim.x & ~1u) && pr < lpg * V; pr += blockDim.x >> 1)
		{
			const int ll = pr / V, c = pr - ll * V;
			if (group * lpg + ll < a.lines)
			{
				const int sc = (group_s + ll * V) % S + c;
				twisted_solve((const float4*)lred + sc, lred[8 * P * S + sc], S, (float*)rv + ll * V + c, lpg * V, P,
							  threadIdx.x & 1);
			}
		}
		__syncthreads();
		if (active)
		{
			const vec u0 = rv[(2 * k) * lpg + line_local];
			const vec u1 = rv[(2 * k + 1) * lpg + line_local];
			// faces touching this line: another axis's face through the whole line, or this axis's faces at
			// the line ends
			bool line_face = false, end_min = false, end_max = false;
			// only lines on a face of another axis and the end blocks of a line can meet a face: the
			// conditions are read for those alone
			bool near_face = start == 0 || start + mk == a.n;
			for (int d = 0; d < g.dims; d++)
				if (d != AXIS)
					near_face |= coord[d] == 0 || coord[d] == grid[d] - 1;
			if (g.fcond && near_face)
			{
				for (int c = 0; c < V; c++)
					for (int d = 0; d < g.dims; d++)
					{
						const bool cmin = g.fcond[(2 * d) * S + s0 + c] != 0;
						const bool cmax = g.fcond[(2 * d + 1) * S + s0 + c] != 0;
						if (d == AXIS)
						{
							end_min |= cmin;
							end_max |= cmax;
						}
						else
							line_face |= (cmin && coord[d] == 0) || (cmax && coord[d] == grid[d] - 1);
					}
			}
			// the block's values are stored in element order, so the stochastic rounding draws its random words
			// from one stream per block (a loop without the face checks for the blocks that meet no face was
			// measured slower: the second copy of the loop costs registers)
			const vec sc = vload<V>(lscale + s0);
			uint32_t h = K == store_f32 ? 0u : sr_seed(off0, key);
			// faces: one per-lane condition for the interior elements (a face of another axis through the line)
			// and one each for the block's first and last element (this axis's faces at the line's ends); waves
			// with no such lane skip the per-element checks
			const bool face_lo = line_face || (start == 0 && end_min);
			const bool face_hi = line_face || (start + mk == a.n && end_max);
			const bool wave_face = __ballot(face_lo || face_hi) != 0;
#pragma unroll
			for (int j = 0; j < M; j++)
			{
				if (has(j))
				{
					vec x;
					if (j == 0)
						x = u0;
					else if (last(j))
						x = u1;
					else
						x = v[j] - vload<V>(ca + j * S) * u0 - vload<V>(cgp + j * S) * u1;
					if (wave_face)
					{
						const bool fj = j == 0 ? (face_lo || (last(j) && face_hi)) : (last(j) ? face_hi : line_face);
						if (fj)
							x = faces_slow<AXIS, V, false, K>(x, g, s0, coord, start + j, sc);
					}
					// NT: FP32 y/z outputs of a large instance are streamed past the caches (measured: y 0.964 ->
					// 0.947 ms, z 1.037 -> 1.023 ms on 400^3 x 8; along x, whose rows stay in cache, and for the
					// small instances of a batch, whose outputs the next sweep finds in cache, they are slower)
					if constexpr (NT)
						__builtin_nontemporal_store(x, (vec*)((float*)g.data + off0 + j * stride));
					else
						dstore_sr_stream<K, V>(g.data, off0 + j * stride, x, h);
				}
			}
		}
		__syncthreads();
	}
}

// sweep_groups for y/z lines whose threads hold V consecutive flattened plane positions that may wrap over
// the substrates (V need not divide S): per-substrate tables in LDS extended periodically to S + V - 1 entries
// and read component by component, per-component substrate and voxel coordinates for the faces.
template <int M, int AXIS, int V, int K, bool FULL = false>
__device__ __forceinline__ void sweep_groups_wrapped(const part_args& a, float* lds, int64_t g_first,
													 int64_t g_end, int64_t g_step)
{
	// this path is held to a tolerance, not to the reference's operation order: multiply-adds may fuse
#pragma clang fp contract(fast)
	using vec = typename vec_of<V>::type;
	const sweep_geometry& g = a.g;
	int S = g.S; // redefined opaquely in every group iteration, below
	const int P = a.P;
	const int lpg = a.lines_per_group;
	const int Sp = S + V - 1; // LDS row length of the per-substrate tables
	const int ncoef = 4 * a.ntypes * a.mb * Sp;
	const int nred = reduced_table_floats(P, S);
	float* lcoef = lds;
	float* lred = lds + ncoef;
	// per-substrate A and storage scale, so the group loop reads no global memory besides the densities
	float* lA = lred + nred;
	float* lscale = lA + Sp;
	vec* rv = (vec*)(lds + part_rows_offset(a.ntypes, a.mb, P, S, Sp)); // [2P][lpg]

	for (int i = threadIdx.x; i < ncoef; i += blockDim.x)
	{
		const int row = i / Sp;
		lcoef[i] = a.coef[row * S + (i - row * Sp) % S];
	}
	for (int i = threadIdx.x; i < nred; i += blockDim.x)
		lred[i] = a.red[i];
	for (int i = threadIdx.x; i < Sp; i += blockDim.x)
	{
		const int si = i % S;
		lA[i] = a.A[si];
		lscale[i] = K == store_f32 ? 1.f : g.scale[si];
	}
	// V table entries from substrate s0 on, component by component (wrapping leaves the start unaligned)
	auto cld = [](const float* p) -> vec {
		vec r;
#pragma unroll
		for (int c = 0; c < V; c++)
			set_comp(r, c, p[c]);
		return r;
	};

	int line_local, k;
	thread_slot<AXIS>(threadIdx.x, 1, P, lpg, line_local, k);
	int64_t stride = line_stride<AXIS>(g); // redefined opaquely in every group iteration, below
	const int grid[3] = { g.X, g.Y, g.Z };
	const size_t fstride = (size_t)a.ntypes * a.mb * Sp;
	const uint32_t key = K == store_f32 ? 0u : rounding_key(g.rounding, 1 + AXIS);

	int start = 0, mk = 0, type = 0;
	if (k >= 0)
	{
		start = k * a.base + min(k, a.rem);
		mk = a.base + (k < a.rem ? 1 : 0);
		type = (P == 1) ? 0 : (k == 0 ? 0 : (k == P - 1 ? 3 : (k < a.rem ? 1 : 2)));
	}
	auto has = [&](int j) { return (FULL && j < M - 1) || j < mk; };
	auto last = [&](int j) { return (!FULL || j >= M - 2) && j == mk - 1; };
	__syncthreads();

	// 16-bit forms load the next group's block while this one is solved (as in sweep_groups)
	constexpr bool prefetch = K != store_f32;
	typename svec_of<K, V>::type pre[prefetch ? M : 1];
	auto fetch = [&](int64_t grp) {
		const int64_t ln = grp * lpg + line_local;
		if (k >= 0 && ln < a.lines)
		{
			int64_t b = 0;
			int sx = 0;
			int cd[3];
			vline_coords<AXIS, V>(a, ln, b, sx, cd);
			const int64_t o = b + (int64_t)start * stride;
#pragma unroll
			for (int j = 0; j < M; j++)
				if (has(j))
					pre[j] = draw<K, V>(g.data, o + j * stride);
		}
	};
	if constexpr (prefetch)
		if (g_first < g_end)
			fetch(g_first);

	for (int64_t group = g_first; group < g_end; group += g_step)
	{
		// The unrolled element loops address memory at j * stride and the LDS tables at j * S for every j; held
		// across the group loop those 16 + 16 uniform offsets overflow the scalar registers and spill into
		// vector-register lanes (hundreds of SGPR spills per kernel).  Marking stride and S as redefined here
		// makes them recomputed per group instead: measured on 640^3 x 16 BF16 (autotuned) sweeps 8.7 / 9.3 /
		// 10.0 -> 7.5 / 8.7 / 8.5 ms, and 10-12 fewer VGPRs in every 4-substrate kernel.
		asm volatile("" : "+s"(stride), "+s"(S));
		const int64_t line = group * lpg + line_local;
		const bool active = k >= 0 && line < a.lines;
		int64_t base = 0;
		int s0 = 0;
		int coord[3] = { 0, 0, 0 };
		if (active)
			vline_coords<AXIS, V>(a, line, base, s0, coord);
		const int64_t off0 = base + (int64_t)start * stride; // element offset of the block's first value
		const float* cr = lcoef + 0 * fstride + (size_t)type * a.mb * Sp + s0;
		const float* cg = lcoef + 1 * fstride + (size_t)type * a.mb * Sp + s0;
		const float* ca = lcoef + 2 * fstride + (size_t)type * a.mb * Sp + s0;
		const float* cgp = lcoef + 3 * fstride + (size_t)type * a.mb * Sp + s0;

		vec v[M];
		if constexpr (prefetch)
		{
			if (active)
			{
#pragma unroll
				for (int j = 0; j < M; j++)
					if (has(j))
						v[j] = dcvt<K, V>(pre[j]);
			}
			if (group + g_step < g_end)
				fetch(group + g_step);
		}
		if (active)
		{
			if constexpr (!prefetch)
			{
#pragma unroll
				for (int j = 0; j < M; j++)
					if (has(j))
						v[j] = dload<K, V>(g.data, off0 + j * stride);
			}
			// 16-bit forms solve on the stored values: the scale is a power of two, so solving on scaled values and
			// scaling the solution give the same bits, and only the Dirichlet values set below need the scale
			// forward elimination inside the block: v_j = r_j (v_j - A v_{j-1}), written as r_j v_j - (r_j A) v_{j-1}
			// so that the recurrence carries one multiply-add
			const vec As = cld(lA + s0);
#pragma unroll
			for (int j = 0; j < M; j++)
			{
				if (has(j))
				{
					const vec r = cld(cr + j * Sp);
					if (j < 2)
						v[j] = r * v[j];
					else
						v[j] = r * v[j] - (r * As) * v[j - 1];
				}
			}
			// backward elimination towards the block's last unknown
#pragma unroll
			for (int j = M - 3; j >= 1; j--)
				if (has(j + 2))
					v[j] = v[j] - cld(cg + j * Sp) * v[j + 1];
			if (has(2))
				v[0] = cld(ca) * (v[0] - cld(cg) * v[1]);
			rv[(2 * k) * lpg + line_local] = v[0];
#pragma unroll
			for (int j = 1; j < M; j++)
				if (last(j))
					rv[(2 * k + 1) * lpg + line_local] = v[j];
		}
		__syncthreads();
		// reduced systems: one pair of adjacent threads per (line, substrate), solving from both ends.  On every
		// axis the first substrate of line l is (l V) mod S (the lines of a row, column or plane position run
		// over the substrate groups), and the group's share of it is uniform across the workgroup.
		const int group_s = (int)((group * lpg * V) % S);
		for (int pr = threadIdx.x >> 1; threadIdx.x < (blockDim.x & ~1u) && pr < lpg * V; pr += blockDim.x >> 1)
		{
			const int ll = pr / V, c = pr - ll * V;
			if (group * lpg + ll < a.lines)
			{
				const int sc = (group_s + ll * V + c) % S;
				twisted_solve((const float4*)lred + sc, lred[8 * P * S + sc], S, (float*)rv + ll * V + c, lpg * V, P,
							  threadIdx.x & 1);
			}
		}
		__syncthreads();
		if (active)
		{
			const vec u0 = rv[(2 * k) * lpg + line_local];
			const vec u1 = rv[(2 * k + 1) * lpg + line_local];
			// faces touching this line: another axis's face through the whole line, or this axis's faces at
			// the line ends
			bool line_face = false, end_min = false, end_max = false;
			// only lines on a face of another axis and the end blocks of a line can meet a face: the
			// conditions are read for those alone
			bool near_face = start == 0 || start + mk == a.n;
			for (int c = 0; c < V; c++)
			{
				int sc, cc[3];
				comp_coords<AXIS, true>(g, coord, s0, c, sc, cc);
				for (int d = 0; d < g.dims; d++)
					if (d != AXIS)
						near_face |= cc[d] == 0 || cc[d] == grid[d] - 1;
			}
			if (g.fcond && near_face)
			{
				for (int c = 0; c < V; c++)
				{
					int sc, cc[3];
					comp_coords<AXIS, true>(g, coord, s0, c, sc, cc);
					for (int d = 0; d < g.dims; d++)
					{
						const bool cmin = g.fcond[(2 * d) * S + sc] != 0;
						const bool cmax = g.fcond[(2 * d + 1) * S + sc] != 0;
						const int cd = cc[d];
						if (d == AXIS)
						{
							end_min |= cmin;
							end_max |= cmax;
						}
						else
							line_face |= (cmin && cd == 0) || (cmax && cd == grid[d] - 1);
					}
				}
			}
			const vec sc = cld(lscale + s0);
			uint32_t h = K == store_f32 ? 0u : sr_seed(off0, key);
			const bool face_lo = line_face || (start == 0 && end_min);
			const bool face_hi = line_face || (start + mk == a.n && end_max);
			const bool wave_face = __ballot(face_lo || face_hi) != 0;
#pragma unroll
			for (int j = 0; j < M; j++)
			{
				if (has(j))
				{
					vec x;
					if (j == 0)
						x = u0;
					else if (last(j))
						x = u1;
					else
						x = v[j] - cld(ca + j * Sp) * u0 - cld(cgp + j * Sp) * u1;
					if (wave_face)
					{
						const bool fj = j == 0 ? (face_lo || (last(j) && face_hi)) : (last(j) ? face_hi : line_face);
						if (fj)
							x = faces_slow<AXIS, V, true, K>(x, g, s0, coord, start + j, sc);
					}
					dstore_sr_stream<K, V>(g.data, off0 + j * stride, x, h);
				}
			}
		}
		__syncthreads();
	}
}

// Waves per SIMD asked of the register allocator: the 4-substrate FP32 kernels and the 16-bit x kernel are
// held to 128 VGPRs (4 waves); measured on 400^3 x 8 with full blocks: FP32 z 1.31 -> 0.99 ms, BF16 x 1.03 ->
// 0.92 ms, while the 16-bit y/z kernels ran slower under a bound of 3 than without one.
template <int V, int K, int AXIS>
constexpr int sweep_waves_per_simd()
{
	return V == 4 && (K == store_f32 || AXIS == 0) ? 4 : 1;
}

// Blocks of M - 1 or M elements (the usual case for long lines) take the FULL form of the group loop.  Measured
// on 400^3 x 8 BF16: y 1.02 -> 0.90 ms, z 1.09 -> 0.98 ms.
template <int M, int AXIS, int V, int K, int MAXT = max_group_threads(V)>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(sweep_waves_per_simd<V, K, AXIS>())))
void sweep_partitioned_kernel(part_args a)
{
	extern __shared__ float lds[];
	constexpr bool PF = K != store_f32 && AXIS != 0, NT = K == store_f32 && AXIS != 0;
	// Workgroups are dispatched to the 8 XCDs in turn, so workgroups b and b + 8 run side by side on one XCD.
	// Where a group's row segment on y/z is shorter than a 128-byte cache line (16-bit groups of fewer than 16
	// lines), they take neighbouring groups, so the line's other part is found in that XCD's L2: measured on
	// 400^3 x 8 BF16 with 8-line groups, DRAM reads of the y/z sweeps 2.05 -> 1.09 GB per launch (z 0.98 ->
	// 0.88 ms).  Elsewhere the plain order is kept (reordered, the FP32 z sweep was 13 % slower).  The
	// reordering is a permutation of the first W - W mod 8 workgroups; the rest keep their index.
	constexpr int e = K == store_f32 ? 4 : 2;
	const int64_t W = gridDim.x, W8 = W - W % 8, b = blockIdx.x;
	const bool reorder = AXIS != 0 && (int64_t)a.lines_per_group * V * e < 128;
	const int64_t first = reorder && b < W8 ? (b % 8) * (W8 / 8) + b / 8 : b;
	if (a.base >= M - 1)
		sweep_groups<M, AXIS, V, K, PF, NT, true>(a, lds, first, a.groups, W);
	else
		sweep_groups<M, AXIS, V, K, PF, NT, false>(a, lds, first, a.groups, W);
}

// Many microenvironments in one launch: workgroups walk a work list of (instance, group range) items.  The
// instances of one launch have V substrates (x) or flattened plane positions (y, z) per thread, with 16-element
// register blocks; on y and z the V positions may cross voxels (substrate vectors wrap).
template <int AXIS, int K, int V>
__global__ __launch_bounds__(V == 1 ? batch_max_threads : 512) void sweep_batched_kernel(
	const part_args* __restrict__ inst, const batch_item* __restrict__ items, int nitems)
{
	extern __shared__ float lds[];
	for (int it = blockIdx.x; it < nitems; it += gridDim.x)
	{
		const batch_item w = items[it];
		const part_args a = inst[w.instance];
		if constexpr (V > 1 && AXIS != 0)
		{
			if (a.base >= batch_chunk - 1)
				sweep_groups_wrapped<batch_chunk, AXIS, V, K, true>(a, lds, w.begin, w.end, 1);
			else
				sweep_groups_wrapped<batch_chunk, AXIS, V, K, false>(a, lds, w.begin, w.end, 1);
		}
		else
		{
			if (a.base >= batch_chunk - 1)
				sweep_groups<batch_chunk, AXIS, V, K, (K != store_f32), false, true>(a, lds, w.begin, w.end, 1);
			else
				sweep_groups<batch_chunk, AXIS, V, K, (K != store_f32), false, false>(a, lds, w.begin, w.end, 1);
		}
	}
}

// ------------------------------------------------------------------------------------------------
// LDS-staged sweep
// ------------------------------------------------------------------------------------------------
//
// Tiles of whole lines move between HBM and LDS as raw bytes with 16-byte accesses wherever the addresses
// allow (any substrate count, any storage form); the partitioned algorithm then runs with one line per
// thread (16-element register blocks) reading and writing the tile in LDS.  Tiles of one axis:
//   x: `span` consecutive rows (z, y), span*X*S contiguous elements; line (row r, s), element i at
//      r*XS + i*S + s;
//   y: `span` consecutive columns q = x*S + s of one z plane, over all y: Y segments of span elements,
//      XS apart in HBM; element i of line q at segment i;
//   z: `span` consecutive positions p of the (y, q) plane, over all z: Z segments, Y*XS apart.
// Each LDS segment starts at the same offset modulo 16 bytes as its HBM source, so the vector part of the
// copy is aligned on both sides; the few elements before the first and after the last 16-byte boundary of a
// segment move one by one.

struct tile_args
{
	part_args a;	 // partition and coefficient tables of the axis (its lines/groups fields are unused)
	int span;		 // x: rows per tile; y/z: columns per tile
	int64_t tiles;	 // tiles of the axis
	int pass_lines;	 // lines solved together (threads = pass_lines * P, rounded up to a wavefront)
	int pitch;		 // y/z: LDS elements per segment (a multiple of 16 bytes, >= span + one vector)
	int64_t tile_elems; // LDS elements of the tile buffer
};

struct tile_view
{
	int64_t g0;		 // HBM element offset of segment 0
	int64_t gstride; // HBM elements between segments
	int nseg, seglen;
	int lines;
	int64_t first; // x: first row; y: z plane; z: first plane position
	int64_t q0;	   // y: first column
};

template <int AXIS>
__device__ __forceinline__ tile_view tile_at(const tile_args& t, int64_t tile)
{
	const sweep_geometry& g = t.a.g;
	const int64_t XS = (int64_t)g.X * g.S;
	tile_view v;
	if constexpr (AXIS == 0)
	{
		const int64_t rows = (int64_t)g.Y * g.Z;
		v.first = tile * t.span;
		const int nrows = (int)min((int64_t)t.span, rows - v.first);
		v.g0 = v.first * XS;
		v.gstride = 0;
		v.nseg = 1;
		v.seglen = (int)(nrows * XS);
		v.lines = nrows * g.S;
		v.q0 = 0;
	}
	else if constexpr (AXIS == 1)
	{
		const int64_t per_plane = (XS + t.span - 1) / t.span;
		v.first = tile / per_plane;
		v.q0 = (tile % per_plane) * t.span;
		v.seglen = (int)min((int64_t)t.span, XS - v.q0);
		v.g0 = v.first * (int64_t)g.Y * XS + v.q0;
		v.gstride = XS;
		v.nseg = g.Y;
		v.lines = v.seglen;
	}
	else
	{
		const int64_t plane = (int64_t)g.Y * XS;
		v.first = tile * t.span;
		v.q0 = 0;
		v.seglen = (int)min((int64_t)t.span, plane - v.first);
		v.g0 = v.first;
		v.gstride = plane;
		v.nseg = g.Z;
		v.lines = v.seglen;
	}
	return v;
}

// LDS element index of segment `sg`'s first element: segments are `pitch` apart (x: one segment at 0),
// shifted by the HBM element offset modulo one 16-byte vector.
template <int AXIS, int VE>
__device__ __forceinline__ int lds_seg(const tile_args& t, const tile_view& v, int sg)
{
	const int shift = (int)((v.g0 + (int64_t)sg * v.gstride) & (VE - 1));
	return (AXIS == 0 ? 0 : sg * t.pitch) + shift;
}

// HBM <-> LDS copy of a tile's raw elements (to_lds: load; else store).
template <int AXIS, int K, bool to_lds>
__device__ __forceinline__ void tile_copy(const tile_args& t, const tile_view& v, typename store_of<K>::elem* tl)
{
	using elem = typename store_of<K>::elem;
	constexpr int VE = 16 / (int)sizeof(elem); // elements per 16-byte vector
	elem* gd = (elem*)t.a.g.data;
	const int maxvec = v.seglen / VE + 1;
	const int64_t nslots = (int64_t)v.nseg * maxvec;
	for (int64_t slot = threadIdx.x; slot < nslots; slot += blockDim.x)
	{
		const int sg = (int)(slot / maxvec);
		const int vi = (int)(slot % maxvec);
		const int64_t gs = v.g0 + (int64_t)sg * v.gstride;
		const int head = (int)((VE - (gs & (VE - 1))) & (VE - 1));
		if (head + (vi + 1) * VE > v.seglen)
			continue;
		const int e = head + vi * VE;
		uint4* gp = (uint4*)(gd + gs + e);
		uint4* lp = (uint4*)(tl + lds_seg<AXIS, VE>(t, v, sg) + e);
		if constexpr (to_lds)
			*lp = *gp;
		else
			*gp = *lp;
	}
	// elements before the first and after the last 16-byte boundary of each segment
	const int64_t nscalar = (int64_t)v.nseg * 2 * (VE - 1);
	for (int64_t slot = threadIdx.x; slot < nscalar; slot += blockDim.x)
	{
		const int sg = (int)(slot / (2 * (VE - 1)));
		const int r = (int)(slot % (2 * (VE - 1)));
		const int64_t gs = v.g0 + (int64_t)sg * v.gstride;
		const int head = (int)min((int64_t)((VE - (gs & (VE - 1))) & (VE - 1)), (int64_t)v.seglen);
		const int body = (v.seglen - head) / VE * VE;
		int e;
		if (r < VE - 1)
		{
			if (r >= head)
				continue;
			e = r;
		}
		else
		{
			e = head + body + (r - (VE - 1));
			if (e >= v.seglen)
				continue;
		}
		elem* gp = gd + gs + e;
		elem* lp = tl + lds_seg<AXIS, VE>(t, v, sg) + e;
		if constexpr (to_lds)
			*lp = *gp;
		else
			*gp = *lp;
	}
}

template <int AXIS, int K>
__device__ __forceinline__ void sweep_tiles(const tile_args& t, float* lds, int64_t first_tile, int64_t end_tile,
											int64_t step)
{
#pragma clang fp contract(fast)
	using elem = typename store_of<K>::elem;
	constexpr int VE = 16 / (int)sizeof(elem);
	constexpr int M = batch_chunk;
	const part_args& a = t.a;
	const sweep_geometry& g = a.g;
	const int S = g.S;
	const int P = a.P;
	const int lpp = t.pass_lines;
	const int ncoef = 4 * a.ntypes * a.mb * S;
	const int nred = reduced_table_floats(P, S);
	float* lcoef = lds;
	float* lred = lds + ncoef;
	float* rv = lred + nred; // [2P][lpp]
	// tile buffer after the tables, 16-byte aligned
	const int off = ((ncoef + nred + 2 * P * lpp) + 3) & ~3;
	elem* tl = (elem*)(lds + off);

	for (int i = threadIdx.x; i < ncoef; i += blockDim.x)
		lcoef[i] = a.coef[i];
	for (int i = threadIdx.x; i < nred; i += blockDim.x)
		lred[i] = a.red[i];

	// thread -> (line in the pass, block k); consecutive lanes take consecutive lines
	const int line_local = threadIdx.x % lpp;
	int k = threadIdx.x / lpp;
	if (k >= P)
		k = -1;
	int start = 0, mk = 0, type = 0;
	if (k >= 0)
	{
		start = k * a.base + min(k, a.rem);
		mk = a.base + (k < a.rem ? 1 : 0);
		type = (P == 1) ? 0 : (k == 0 ? 0 : (k == P - 1 ? 3 : (k < a.rem ? 1 : 2)));
	}
	const size_t fstride = (size_t)a.ntypes * a.mb * S;
	const uint32_t key = K == store_f32 ? 0u : rounding_key(g.rounding, 1 + AXIS);
	const int grid[3] = { g.X, g.Y, g.Z };
	const int64_t XS = (int64_t)g.X * S;
	__syncthreads();

	for (int64_t tile = first_tile; tile < end_tile; tile += step)
	{
		const tile_view v = tile_at<AXIS>(t, tile);
		tile_copy<AXIS, K, true>(t, v, tl);
		__syncthreads();
		for (int l0 = 0; l0 < v.lines; l0 += lpp)
		{
			const int l = l0 + line_local;
			const bool active = k >= 0 && l < v.lines;
			// the line: its substrate, voxel coordinates and LDS addressing
			int s = 0, lbase = 0, lstride = 0;
			int coord[3] = { 0, 0, 0 };
			int64_t gbase = 0, gstep = 0; // HBM element index of element i: gbase + i * gstep
			if (active)
			{
				if constexpr (AXIS == 0)
				{
					const int r = l / S;
					s = l - r * S;
					const int64_t row = v.first + r;
					coord[1] = (int)(row % g.Y);
					coord[2] = (int)(row / g.Y);
					lbase = lds_seg<AXIS, VE>(t, v, 0) + r * (int)XS + s;
					lstride = S;
					gbase = v.g0 + (int64_t)r * XS + s;
					gstep = S;
				}
				else
				{
					const int64_t q = (AXIS == 1 ? v.q0 : v.first) + l; // column (y) or plane position (z)
					s = (int)(q % S);
					coord[0] = (int)((q / S) % g.X);
					if constexpr (AXIS == 1)
						coord[2] = (int)v.first;
					else
						coord[1] = (int)(q / XS);
					lbase = l;
					lstride = t.pitch;
					gbase = v.g0 + l;
					gstep = v.gstride;
				}
			}
			// element i of this line in LDS (y/z: segment i is shifted by its HBM offset modulo a vector)
			const int sh0 = (int)(v.g0 & (VE - 1)), shs = (int)(v.gstride & (VE - 1));
			auto at = [&](int i) -> elem& {
				if constexpr (AXIS == 0)
					return tl[lbase + i * lstride];
				else
					return tl[lbase + i * lstride + ((sh0 + i * shs) & (VE - 1))];
			};
			const float* cr = lcoef + 0 * fstride + (size_t)type * a.mb * S + s;
			const float* cg = lcoef + 1 * fstride + (size_t)type * a.mb * S + s;
			const float* ca = lcoef + 2 * fstride + (size_t)type * a.mb * S + s;
			const float* cgp = lcoef + 3 * fstride + (size_t)type * a.mb * S + s;

			float x[M];
			if (active)
			{
				// 16-bit forms solve on the stored (power-of-two scaled) values, as in sweep_groups
#pragma unroll
				for (int j = 0; j < M; j++)
					if (j < mk)
						x[j] = (float)at(start + j);
				const float As = a.A[s];
#pragma unroll
				for (int j = 0; j < M; j++)
				{
					if (j < mk)
					{
						const float r = cr[j * S];
						if (j < 2)
							x[j] = r * x[j];
						else
							x[j] = r * x[j] - (r * As) * x[j - 1];
					}
				}
#pragma unroll
				for (int j = M - 3; j >= 1; j--)
					if (j <= mk - 3)
						x[j] = x[j] - cg[j * S] * x[j + 1];
				if (mk >= 3)
					x[0] = ca[0] * (x[0] - cg[0] * x[1]);
				rv[(2 * k) * lpp + line_local] = x[0];
#pragma unroll
				for (int j = 1; j < M; j++)
					if (j == mk - 1)
						rv[(2 * k + 1) * lpp + line_local] = x[j];
			}
			__syncthreads();
			// reduced systems: one pair of adjacent threads per line of the pass, solving from both ends
			for (int pr = threadIdx.x >> 1; threadIdx.x < (blockDim.x & ~1u) && pr < lpp; pr += blockDim.x >> 1)
			{
				const int pl = l0 + pr;
				if (pl < v.lines)
				{
					const int ps = AXIS == 0 ? pl % S : (int)(((AXIS == 1 ? v.q0 : v.first) + pl) % S);
					twisted_solve((const float4*)lred + ps, lred[8 * P * S + ps], S, rv + pr, lpp, P, threadIdx.x & 1);
				}
			}
			__syncthreads();
			if (active)
			{
				const float u0 = rv[(2 * k) * lpp + line_local];
				const float u1 = rv[(2 * k + 1) * lpp + line_local];
				face_epilogue fe;
				fe.init(g.fval, g.fcond, S, s, AXIS, coord, grid, g.dims, K == store_f32 ? 1.f : g.scale[s]);
#pragma unroll
				for (int j = 0; j < M; j++)
				{
					if (j < mk)
					{
						float y;
						if (j == 0)
							y = u0;
						else if (j == mk - 1)
							y = u1;
						else
							y = x[j] - ca[j * S] * u0 - cgp[j * S] * u1;
						const int i = start + j;
						y = fe.apply(y, i, a.n);
						if constexpr (K == store_f32)
							at(i) = y;
						else
						{
							// stochastic rounding keyed by the HBM element index (independent of the tiling)
							const int64_t gi = gbase + (int64_t)i * gstep;
							constexpr uint32_t drop = K == store_bf16 ? 0xFFFFu : 0x1FFFu;
							const uint32_t rb = mix32((uint32_t)gi ^ ((uint32_t)((uint64_t)gi >> 32) * 0x9E3779B9u) ^ key);
							const uint32_t b = (__float_as_uint(y) + (rb & drop)) & ~drop;
							at(i) = (elem)__uint_as_float(b);
						}
					}
				}
			}
			__syncthreads();
		}
		tile_copy<AXIS, K, false>(t, v, tl);
		__syncthreads();
	}
}

template <int AXIS, int K>
__global__ __launch_bounds__(1024) void sweep_staged_kernel(tile_args t)
{
	extern __shared__ float lds[];
	sweep_tiles<AXIS, K>(t, lds, blockIdx.x, t.tiles, gridDim.x);
}

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------

static constexpr int chunk_sizes[] = { 8, 16, 32, 48, 64 };

int supported_chunk(int request)
{
	for (int c : chunk_sizes)
		if (request <= c)
			return c;
	return 64;
}

static sweep_geometry geometry(bfvm_ctx* ctx)
{
	sweep_geometry g;
	g.data = ctx->dens;
	g.scale = ctx->dscale;
	g.inv = ctx->dinv;
	g.rounding = ctx->rounding_step;
	g.X = ctx->mesh.grid[0];
	g.Y = ctx->mesh.grid[1];
	g.Z = ctx->mesh.grid[2];
	g.S = ctx->mesh.S;
	g.dims = ctx->mesh.dims;
	g.fval = ctx->has_faces ? ctx->face_val : nullptr;
	g.fcond = ctx->has_faces ? ctx->face_cond : nullptr;
	return g;
}

void launch_exact_coefficients(bfvm_ctx* ctx, int axis)
{
	auto& t = ctx->axes[axis];
	const int S = ctx->mesh.S;
	hipLaunchKernelGGL(exact_coefficients_kernel, dim3((S + 63) / 64), dim3(64), 0, ctx->stream, t.exact_b, t.exact_c,
					   t.exact_e, ctx->d_D, ctx->d_decay, ctx->mesh.dt, ctx->mesh.vs[axis], ctx->mesh.dims,
					   ctx->mesh.grid[axis], S);
	BFVM_HIP(hipGetLastError());
}

static void free_partition_tables(axis_tables& t)
{
	if (t.coef)
		BFVM_HIP(hipFree(t.coef));
	if (t.red)
		BFVM_HIP(hipFree(t.red));
	if (t.A)
		BFVM_HIP(hipFree(t.A));
	t.coef = t.red = t.A = nullptr;
}

// Most threads of a register-sweep workgroup: 4-substrate sweeps without the 16-bit prefetch fit 1024 threads
// (124-128 VGPRs), with it they keep 512 (160 VGPRs); one substrate per thread always fits 1024.
static int partition_max_threads(int vec, int storage, int axis)
{
	if (vec != 4)
		return 1024;
	return (storage == BFVM_STORAGE_F32 || axis == 0) ? 1024 : 512;
}

// Lines handled by one workgroup and its thread count for a partition of P blocks.  `request` is the
// number of lines per workgroup for y/z and the number of x rows per workgroup for x; 0 picks the shapes
// measured fastest on 400^3 x 8 (x: as many rows as fit in 200 threads; y: 8 lines of 4-substrate vectors,
// z: 16; one substrate per thread: 32).
static void group_shape(int axis, int SG, int P, int request, int max_threads, int& lpg, int& threads,
						bool vec4 = false)
{
	if (axis == 0)
	{
		const int per_row = SG * P; // SG = substrate groups of V per x row
		const int rows = request > 0 ? request : std::max(1, 200 / per_row);
		lpg = rows * SG;
		threads = ((rows * per_row + 63) / 64) * 64;
	}
	else
	{
		lpg = request > 0 ? request : (!vec4 && max_threads >= 1024 ? 32 : axis == 1 ? 8 : 16);
		if (lpg * P > max_threads)
			lpg = std::max(8, max_threads / P / 8 * 8);
		threads = ((lpg * P + 63) / 64) * 64;
	}
}

static size_t partition_lds_bytes(const partition& p, int S, int lpg, bool wrap = false)
{
	const int Sp = wrap ? S + p.vec - 1 : S;
	return sizeof(float) * ((size_t)part_rows_offset(p.ntypes, p.mb, p.P, S, Sp) + (size_t)2 * p.P * lpg * p.vec);
}

partition choose_partition(const bfvm_ctx* ctx, int axis, int chunk, int lines_request, int max_vec)
{
	const int n = ctx->mesh.grid[axis];
	const int S = ctx->mesh.S;
	partition none;
	none.n = n;
	if (n <= 1)
		return none;
	// 4 substrates per thread when they pack into 16-byte vectors (register chunk 16: 64 data VGPRs)
	const bool vec4 = max_vec >= 4 && S % 4 == 0;
	auto vec_for = [&](int M) { return (vec4 && M == 16) ? 4 : 1; };
	// candidate register chunks, smallest first from the request; 16-bit storage instantiates 16 and 32 only
	const bool wide = ctx->storage != BFVM_STORAGE_F32;
	std::vector<int> candidates;
	const int first = chunk > 0 ? chunk : (vec4 ? 16 : 32);
	for (int c : chunk_sizes)
		if (c >= supported_chunk(first) && (!wide || c == 16 || c == 32))
			candidates.push_back(c);
	if (candidates.empty())
		candidates.push_back(32);
	for (size_t ci = 0;; ci++)
	{
		const int M = candidates[ci];
		const bool last_candidate = ci + 1 == candidates.size();
		partition p;
		p.n = n;
		p.chunk = M;
		p.vec = vec_for(M);
		p.P = (n + M - 1) / M;
		p.base = n / p.P;
		p.rem = n % p.P;
		p.ntypes = p.P == 1 ? 1 : 4;
		p.mb = p.P == 1 ? n : p.base + 1;
		int lpg, threads;
		// staged and batched sweeps (max_vec 1) keep their own limits; the register sweeps may widen (see
		// partition_max_threads)
		const int max_threads = max_vec >= 4 ? partition_max_threads(p.vec, ctx->storage, axis)
											 : (p.vec == 4 ? 512 : 1024);
		group_shape(axis, S / p.vec, p.P, lines_request, max_threads, lpg, threads, p.vec == 4);
		p.lpg = lpg;
		p.threads = threads;
		const bool fits = threads <= max_threads && lpg * p.P <= threads &&
						  partition_lds_bytes(p, S, lpg) <= 64 * 1024 && (p.P == 1 || p.base >= 3);
		if (fits)
			return p;
		if (last_candidate)
			throw std::runtime_error("partitioned sweep: axis too long for the register/LDS budget "
									 "(use the exact sweep)");
	}
}

void alloc_partition_tables(const partition& p, int S, float*& coef, float*& red, float*& A)
{
	BFVM_HIP(hipMalloc(&coef, sizeof(float) * 4 * p.ntypes * p.mb * S));
	BFVM_HIP(hipMalloc(&red, sizeof(float) * reduced_table_floats(p.P, S)));
	BFVM_HIP(hipMalloc(&A, sizeof(float) * S));
}

void build_partition(bfvm_ctx* ctx, int axis, int chunk, int lines_request)
{
	auto& t = ctx->axes[axis];
	free_partition_tables(t);
	// the staged sweep runs one line per thread on 16-element blocks
	const bool staged = ctx->algorithm == BFVM_SWEEP_STAGED;
	// 16-bit forms take z lines in groups of 8, not 16 (measured on 400^3 x 8 BF16: 1.38 -> 1.11 ms)
	if (lines_request == 0 && axis == 2 && ctx->storage != BFVM_STORAGE_F32 && !staged)
		lines_request = 8;
	t.part = choose_partition(ctx, axis, staged ? batch_chunk : chunk, lines_request, staged ? 1 : 4);
	if (t.part.P > 0)
		alloc_partition_tables(t.part, ctx->mesh.S, t.coef, t.red, t.A);
}

void launch_partition_tables(const bfvm_ctx* ctx, int axis, const partition& p, float* coef, float* red, float* A,
							 hipStream_t stream)
{
	if (p.P == 0)
		return;
	const int S = ctx->mesh.S;
	hipLaunchKernelGGL(partition_coefficients_kernel, dim3((S + 63) / 64), dim3(64), 0, stream, coef, red, A,
					   ctx->d_D, ctx->d_decay, ctx->mesh.dt, ctx->mesh.vs[axis], ctx->mesh.dims, p.n, S, p.P, p.base,
					   p.rem, p.ntypes, p.mb);
	BFVM_HIP(hipGetLastError());
}

void launch_partition_coefficients(bfvm_ctx* ctx, int axis)
{
	auto& t = ctx->axes[axis];
	launch_partition_tables(ctx, axis, t.part, t.coef, t.red, t.A, ctx->stream);
}

// Workgroups of a single-instance sweep; each loops over groups when there are more.
#ifndef BFVM_SWEEP_MAX_BLOCKS
#define BFVM_SWEEP_MAX_BLOCKS 8192
#endif
static constexpr int64_t sweep_max_blocks = BFVM_SWEEP_MAX_BLOCKS;

// (mul, shift) of fdiv for divisor d (1 <= d < 2^31)
static void fast_divisor(uint32_t d, uint32_t& mul, int& shift)
{
	int l = 0;
	while ((1ull << l) < d)
		l++;
	mul = (uint32_t)(((1ull << (31 + l)) + d - 1) / d);
	shift = 31 + l;
}

static part_args make_part_args(int axis, const sweep_geometry& g, const partition& p, const float* coef,
								const float* red, const float* A)
{
	part_args a;
	a.g = g;
	a.n = p.n;
	a.P = p.P;
	a.base = p.base;
	a.rem = p.rem;
	a.ntypes = p.ntypes;
	a.mb = p.mb;
	a.coef = coef;
	a.red = red;
	a.A = A;
	const int64_t total = (int64_t)g.X * g.Y * g.Z * g.S;
	a.lines = total / ((int64_t)p.n * p.vec);
	a.vec = p.vec;
	a.lines_per_group = p.lpg;
	a.groups = (a.lines + p.lpg - 1) / p.lpg;
	if ((uint64_t)a.lines * p.vec >= (1ull << 31))
		throw std::runtime_error("sweep: more than 2^31 line positions along one axis (mesh too large)");
	const int64_t XS = (int64_t)g.X * g.S;
	const int64_t d0 = axis == 0 ? g.S / p.vec : (axis == 1 ? XS / p.vec : g.S);
	const int64_t d1 = axis == 0 ? g.Y : (axis == 1 ? g.S : g.X);
	fast_divisor((uint32_t)d0, a.div_mul[0], a.div_shift[0]);
	fast_divisor((uint32_t)d1, a.div_mul[1], a.div_shift[1]);
	return a;
}

// the 1024-thread instantiation exists where partition_max_threads allows it
template <int V, int K, int AXIS>
constexpr bool wide_group_ok() { return V == 4 && (K == store_f32 || AXIS == 0); }

template <int M, int AXIS, int V, int K>
static void launch_partitioned_t(bfvm_ctx* ctx, const sweep_geometry& g, const axis_tables& t)
{
	const partition& p = t.part;
	const part_args a = make_part_args(AXIS, g, p, t.coef, t.red, t.A);
	const int threads = p.threads;
	const size_t lds = partition_lds_bytes(p, g.S, p.lpg);
	const int64_t cap = ctx->sweep_workgroups[AXIS] > 0 ? ctx->sweep_workgroups[AXIS] : sweep_max_blocks;
	const int blocks = (int)std::min<int64_t>(a.groups, cap);
	if constexpr (wide_group_ok<V, K, AXIS>())
	{
		if (threads > max_group_threads(V))
		{
			hipLaunchKernelGGL((sweep_partitioned_kernel<M, AXIS, V, K, 1024>), dim3(blocks), dim3(threads), lds,
							   ctx->stream, a);
			BFVM_HIP(hipGetLastError());
			return;
		}
	}
	if (threads > max_group_threads(V))
		throw std::logic_error("partitioned sweep: workgroup larger than the kernel allows");
	hipLaunchKernelGGL((sweep_partitioned_kernel<M, AXIS, V, K>), dim3(blocks), dim3(threads), lds, ctx->stream, a);
	BFVM_HIP(hipGetLastError());
}

// FP32 storage: every register chunk; 16-bit storage: chunks 16 and 32 (choose_partition restricts them)
template <int AXIS, int K>
static void launch_partitioned_axis(bfvm_ctx* ctx, const sweep_geometry& g, const axis_tables& t)
{
	if (t.part.vec == 4)
	{
		launch_partitioned_t<16, AXIS, 4, K>(ctx, g, t);
		return;
	}
	if constexpr (K != store_f32)
	{
		if (t.part.chunk == 16)
			launch_partitioned_t<16, AXIS, 1, K>(ctx, g, t);
		else
			launch_partitioned_t<32, AXIS, 1, K>(ctx, g, t);
		return;
	}
	else
	{
		switch (t.part.chunk)
		{
		case 8:
			launch_partitioned_t<8, AXIS, 1, K>(ctx, g, t);
			break;
		case 16:
			launch_partitioned_t<16, AXIS, 1, K>(ctx, g, t);
			break;
		case 32:
			launch_partitioned_t<32, AXIS, 1, K>(ctx, g, t);
			break;
		case 48:
			launch_partitioned_t<48, AXIS, 1, K>(ctx, g, t);
			break;
		default:
			launch_partitioned_t<64, AXIS, 1, K>(ctx, g, t);
			break;
		}
	}
}

template <int AXIS>
static void launch_exact_axis(bfvm_ctx* ctx, const sweep_geometry& g, const axis_tables& t)
{
	const int64_t total = (int64_t)g.X * g.Y * g.Z * g.S;
	const int n = ctx->mesh.grid[AXIS];
	const int64_t lines = total / n;
	const int64_t blocks = (lines + 255) / 256;
	hipLaunchKernelGGL((sweep_exact_kernel<AXIS>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream, g, lines, n,
					   t.exact_b, t.exact_c, t.exact_e);
	BFVM_HIP(hipGetLastError());
}

// A workgroup's LDS holds the coefficient tables, the reduced rows of one pass and the tile; the whole
// stays within 64 KB (tiles of at most 48 KB); one pass solves up to 1024 / P lines.
static constexpr size_t staged_lds_limit = 64 * 1024;
static constexpr size_t staged_tile_bytes = 48 * 1024;

static tile_args make_tile_args(const bfvm_ctx* ctx, int axis, const sweep_geometry& g, const partition& p,
								const float* coef, const float* red, const float* A)
{
	tile_args t;
	t.a = make_part_args(axis, g, p, coef, red, A);
	const size_t e = storage_bytes(ctx->storage);
	const int VE = (int)(16 / e);
	const int64_t XS = (int64_t)g.X * g.S;
	const int n = p.n;
	const int max_lines = std::max(1, 1024 / p.P);
	const size_t tables = sizeof(float) * ((size_t)4 * p.ntypes * p.mb * g.S + (size_t)reduced_table_floats(p.P, g.S) +
										   (size_t)2 * p.P * max_lines + 4);
	if (tables + 16 * n * e > staged_lds_limit)
		throw std::runtime_error("staged sweep: coefficient tables leave no room for a tile");
	const size_t budget = std::min(staged_tile_bytes, staged_lds_limit - tables);
	if (axis == 0)
	{
		const int64_t rows = (int64_t)g.Y * g.Z;
		t.span = (int)std::max<int64_t>(1, std::min<int64_t>(rows, (budget - 16) / (XS * e)));
		t.tiles = (rows + t.span - 1) / t.span;
		t.pitch = 0;
		t.tile_elems = t.span * XS + VE;
		t.pass_lines = (int)std::min<int64_t>(t.span * (int64_t)g.S, max_lines);
	}
	else
	{
		const int64_t width = axis == 1 ? XS : (int64_t)g.Y * XS; // columns of a plane / plane positions
		// pitch = span rounded up to a vector, plus one vector of shift room
		int span = (int)(budget / ((size_t)n * e)) / VE * VE - VE;
		span = std::max(VE, std::min(span, 256));
		t.span = (int)std::min<int64_t>(span, (width + VE - 1) / VE * VE);
		t.pitch = (t.span + VE - 1) / VE * VE + VE;
		t.tiles = (axis == 1 ? (int64_t)g.Z : 1) * ((width + t.span - 1) / t.span);
		t.tile_elems = (int64_t)n * t.pitch;
		t.pass_lines = std::min(t.span, max_lines);
	}
	return t;
}

static size_t tile_lds_bytes(const tile_args& t, int S, int storage)
{
	const int P = t.a.P;
	const size_t tables =
		(size_t)4 * t.a.ntypes * t.a.mb * S + (size_t)reduced_table_floats(P, S) + (size_t)2 * P * t.pass_lines;
	return ((tables + 3) & ~(size_t)3) * sizeof(float) + (size_t)t.tile_elems * storage_bytes(storage);
}

static int tile_threads(const tile_args& t) { return (t.pass_lines * t.a.P + 63) / 64 * 64; }

template <int AXIS, int K>
static void launch_staged_axis(bfvm_ctx* ctx, const sweep_geometry& g, const axis_tables& t)
{
	if (t.part.chunk != batch_chunk)
		throw std::logic_error("staged sweep needs the 16-element partition");
	const tile_args ta = make_tile_args(ctx, AXIS, g, t.part, t.coef, t.red, t.A);
	const size_t lds = tile_lds_bytes(ta, g.S, ctx->storage);
	if (lds > staged_lds_limit)
		throw std::runtime_error("staged sweep: tile does not fit LDS");
	const int blocks =
		(int)std::min<int64_t>(ta.tiles, ctx->sweep_workgroups[AXIS] > 0 ? ctx->sweep_workgroups[AXIS] : 4096);
	hipLaunchKernelGGL((sweep_staged_kernel<AXIS, K>), dim3(blocks), dim3(tile_threads(ta)), lds, ctx->stream, ta);
	BFVM_HIP(hipGetLastError());
}

template <int K>
static void launch_sweep_storage(bfvm_ctx* ctx, int axis)
{
	const sweep_geometry g = geometry(ctx);
	const axis_tables& t = ctx->axes[axis];
	const int n = ctx->mesh.grid[axis];
	if (n == 1)
	{
		const int64_t total = (int64_t)g.X * g.Y * g.Z * g.S;
		hipLaunchKernelGGL((sweep_single_kernel<K>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream,
						   g, axis, t.exact_b);
		BFVM_HIP(hipGetLastError());
		return;
	}
	if (ctx->algorithm == BFVM_SWEEP_EXACT)
	{
		if constexpr (K != store_f32)
			throw std::runtime_error("the exact sweep needs FP32 density storage");
		else if (axis == 0)
			launch_exact_axis<0>(ctx, g, t);
		else if (axis == 1)
			launch_exact_axis<1>(ctx, g, t);
		else
			launch_exact_axis<2>(ctx, g, t);
		return;
	}
	if (ctx->algorithm == BFVM_SWEEP_STAGED)
	{
		if (axis == 0)
			launch_staged_axis<0, K>(ctx, g, t);
		else if (axis == 1)
			launch_staged_axis<1, K>(ctx, g, t);
		else
			launch_staged_axis<2, K>(ctx, g, t);
		return;
	}
	if (axis == 0)
		launch_partitioned_axis<0, K>(ctx, g, t);
	else if (axis == 1)
		launch_partitioned_axis<1, K>(ctx, g, t);
	else
		launch_partitioned_axis<2, K>(ctx, g, t);
}

void launch_sweep(bfvm_ctx* ctx, int axis)
{
	switch (ctx->storage)
	{
	case BFVM_STORAGE_BF16:
		launch_sweep_storage<store_bf16>(ctx, axis);
		break;
	case BFVM_STORAGE_F16:
		launch_sweep_storage<store_f16>(ctx, axis);
		break;
	default:
		launch_sweep_storage<store_f32>(ctx, axis);
		break;
	}
}

} // namespace bfvm

namespace bfvm {

// ------------------------------------------------------------------------------------------------
// batched sweeps
// ------------------------------------------------------------------------------------------------

void free_batch_sweep(bfvm_batch* b, int axis)
{
	auto& ax = b->axes[axis];
	for (float* t : ax.tables)
		if (t)
			BFVM_HIP(hipFree(t));
	ax.tables.clear();
	for (auto& part : ax.parts)
	{
		if (part.inst)
			BFVM_HIP(hipFree(part.inst));
		if (part.items)
			BFVM_HIP(hipFree(part.items));
	}
	ax = bfvm_batch::sweep_plan {};
}

// Per instance: the partition with 16-element register blocks; its own coefficient tables; work items of about
// 64K elements.  Each thread holds V = 4 or 1 consecutive elements of a row: on x, V substrates of a voxel
// (V divides S); on y and z, V flattened plane positions (V divides the row or plane length, so every vector
// is aligned), which wrap over the substrates when V does not divide S.  4-element vectors move 16 (FP32) or
// 8 bytes per lane, against 4 or 2 for one element.  (Two-element vectors made the compiler move the register
// blocks through scratch memory, so they are not used.)
static int batch_vector_width(const bfvm_ctx* ctx, int axis)
{
	const int S = ctx->mesh.S;
	const int64_t XS = (int64_t)ctx->mesh.grid[0] * S;
	const int64_t unit = axis == 0 ? S : (axis == 1 ? XS : (int64_t)ctx->mesh.grid[1] * XS);
	return unit % 4 == 0 ? 4 : 1;
}

void prepare_batch_sweep(bfvm_batch* b, int axis)
{
	auto& ax = b->axes[axis];
	std::vector<part_args> args[2];
	std::vector<batch_item> items[2];
	for (auto& a : args)
		a.assign(b->ctx.size(), part_args {});
	for (size_t i = 0; i < b->ctx.size(); i++)
	{
		bfvm_ctx* ctx = b->ctx[i];
		if (axis >= ctx->mesh.dims)
			continue;
		const int n = ctx->mesh.grid[axis];
		if (n == 1)
		{
			ax.single.push_back((int)i);
			continue;
		}
		const int S = ctx->mesh.S;
		const int V = batch_vector_width(ctx, axis);
		const int w = V == 4 ? 0 : 1;
		const int max_threads = V == 1 ? batch_max_threads : 512;
		partition p = choose_partition(ctx, axis, batch_chunk, 1, 1);
		if (p.chunk != batch_chunk)
			throw std::runtime_error("batched sweep: instance " + std::to_string(i) +
									 " does not fit 16-element register blocks (axis too long)");
		p.vec = V;
		const int SG = axis == 0 ? S / V : 1;
		const int target = b->tuning.group_threads[axis][w];
		const int request = axis == 0 ? std::max(1, target / (SG * p.P)) : std::max(8, target / p.P / 8 * 8);
		group_shape(axis, SG, p.P, request, max_threads, p.lpg, p.threads);
		if (p.threads > max_threads)
			throw std::runtime_error("batched sweep: instance " + std::to_string(i) +
									 " does not fit one workgroup (axis too long)");
		float *coef = nullptr, *red = nullptr, *A = nullptr;
		alloc_partition_tables(p, S, coef, red, A);
		ax.tables.push_back(coef);
		ax.tables.push_back(red);
		ax.tables.push_back(A);
		launch_partition_tables(ctx, axis, p, coef, red, A, b->stream);
		sweep_geometry g = geometry(ctx);
		g.rounding = b->rounding;
		args[w][i] = make_part_args(axis, g, p, coef, red, A);
		auto& part = ax.parts[w];
		part.lds = std::max(part.lds, partition_lds_bytes(p, S, p.lpg, V > 1 && axis != 0));
		part.threads = std::max(part.threads, p.threads);
		const int64_t per_group = (int64_t)p.lpg * n * V;
		const int64_t per_item = std::max<int64_t>(1, 65536 / per_group);
		for (int64_t g0 = 0; g0 < args[w][i].groups; g0 += per_item)
			items[w].push_back(batch_item { (int32_t)i, 0, g0, std::min(args[w][i].groups, g0 + per_item) });
		ax.any = true;
	}
	for (int w = 0; w < 2; w++)
	{
		auto& part = ax.parts[w];
		if (items[w].empty())
			continue;
		if (part.lds > 64 * 1024)
			throw std::runtime_error("batched sweep: coefficient tables exceed 64 KB of LDS");
		BFVM_HIP(hipMalloc(&part.inst, sizeof(part_args) * args[w].size()));
		BFVM_HIP(hipMemcpy(part.inst, args[w].data(), sizeof(part_args) * args[w].size(), hipMemcpyHostToDevice));
		BFVM_HIP(hipMalloc(&part.items, sizeof(batch_item) * items[w].size()));
		BFVM_HIP(hipMemcpy(part.items, items[w].data(), sizeof(batch_item) * items[w].size(), hipMemcpyHostToDevice));
		part.nitems = (int)items[w].size();
		part.blocks = std::min(part.nitems, b->tuning.max_blocks[axis][w]);
	}
}

template <int AXIS, int K>
static void launch_batched_t(bfvm_batch* b)
{
	const auto& ax = b->axes[AXIS];
	const auto& p4 = ax.parts[0];
	const auto& p1 = ax.parts[1];
	if (p4.nitems)
		hipLaunchKernelGGL((sweep_batched_kernel<AXIS, K, 4>), dim3(p4.blocks), dim3(p4.threads), p4.lds, b->stream,
						   (const part_args*)p4.inst, p4.items, p4.nitems);
	if (p1.nitems)
		hipLaunchKernelGGL((sweep_batched_kernel<AXIS, K, 1>), dim3(p1.blocks), dim3(p1.threads), p1.lds, b->stream,
						   (const part_args*)p1.inst, p1.items, p1.nitems);
	BFVM_HIP(hipGetLastError());
}

void launch_batched_sweep(bfvm_batch* b, int axis)
{
	if (!b->axes[axis].any)
		return;
	by_storage(b->storage, [&](auto kind) {
		constexpr int K = decltype(kind)::value;
		if (axis == 0)
			launch_batched_t<0, K>(b);
		else if (axis == 1)
			launch_batched_t<1, K>(b);
		else
			launch_batched_t<2, K>(b);
	});
}

} // namespace bfvm
