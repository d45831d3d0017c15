/*
 * TEST INFRASTRUCTURE ONLY — CPU oracle for paraBioFVM's host solver path.
 * See biofvm_oracle.h for what it is, how it is built and what pins it.
 *
 * The arithmetic of every expression follows the reference's host solver
 * expression-for-expression (same operand order, same types, no contraction),
 * so the float build reproduces the reference's float results.  The OpenMP
 * structure (one parallel region, orphaned worksharing loops) follows the
 * reference's host path so the oracle doubles as the CPU baseline.
 *
 * Behaviour defined here where the reference has none:
 *   - an agent whose position maps outside the grid (or is not finite) is ignored
 *     (the reference indexes out of bounds: src/mesh.cpp:47-53 does not clamp);
 *   - a gradient along an axis of length < 2 is rejected (the reference reads
 *     past the line: host/gradient_solver.cpp:25-27).
 */
#include "biofvm_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct or_state
{
	int dims;
	int mins[3], maxs[3], vs[3], grid[3];
	int S;
	or_real dt;
	or_real* D;
	or_real* decay;

	or_real* dens;
	or_real* grad;

	or_real* bval[2][3];
	unsigned char* bcond[2][3];

	int n_interior;
	int* ivox;
	or_real* ival;
	unsigned char* icond;

	int A;
	const or_real *sec, *sat, *upt, *nex, *frac, *vol, *pos;
	or_real* intern;
	int compute_internalized;

	/* diffusion: b[i][copies*S], e[i][copies*S], c[copies*S] per axis (host/diffusion_solver.h:398-400) */
	or_real *b[3], *c[3], *e[3];
	int copies[3];
	int substrate_factor;

	/* cell solver (host/cell_solver.h:424-436) */
	or_real *num, *den, *fac;
	size_t cap;
	or_real *rnum, *rden, *rfac;
	int* ballots;
	int is_conflict;
	int initialized;
};

static __thread char g_err[512];

static int fail(const char* msg)
{
	snprintf(g_err, sizeof(g_err), "%s", msg);
	return -1;
}

const char* OR_FN(or_last_error)(void) { return g_err; }

static size_t voxel_count(const or_state* st) { return (size_t)st->grid[0] * st->grid[1] * st->grid[2]; }

/* mesh.cpp:7-25 */
or_state* OR_FN(or_create)(int dims, const int* mins, const int* maxs, const int* voxel_shape, int substrates,
						   or_real time_step, const or_real* diffusion_coefficients, const or_real* decay_rates,
						   const or_real* initial_conditions)
{
	if (dims < 1 || dims > 3 || substrates < 1)
	{
		fail("or_create: bad dims or substrate count");
		return NULL;
	}
	or_state* st = (or_state*)calloc(1, sizeof(or_state));
	st->dims = dims;
	for (int d = 0; d < 3; d++)
	{
		st->mins[d] = mins[d];
		st->maxs[d] = maxs[d];
		st->vs[d] = voxel_shape[d];
		st->grid[d] = 1;
	}
	for (int d = 0; d < dims; d++)
		st->grid[d] = (maxs[d] - mins[d] + voxel_shape[d] - 1) / voxel_shape[d];
	st->S = substrates;
	st->dt = time_step;
	st->D = (or_real*)malloc(sizeof(or_real) * substrates);
	st->decay = (or_real*)malloc(sizeof(or_real) * substrates);
	memcpy(st->D, diffusion_coefficients, sizeof(or_real) * substrates);
	memcpy(st->decay, decay_rates, sizeof(or_real) * substrates);

	size_t n = voxel_count(st) * (size_t)substrates;
	st->dens = (or_real*)malloc(sizeof(or_real) * n);
	st->grad = (or_real*)calloc(n * (size_t)dims, sizeof(or_real));
	/* microenvironment.cpp:14-25 — fill with the per-substrate initial condition */
#pragma omp parallel for
	for (long long v = 0; v < (long long)voxel_count(st); v++)
		for (int s = 0; s < substrates; s++)
			st->dens[(size_t)v * substrates + s] = initial_conditions[s];
	return st;
}

void OR_FN(or_destroy)(or_state* st)
{
	if (!st)
		return;
	free(st->D);
	free(st->decay);
	free(st->dens);
	free(st->grad);
	for (int m = 0; m < 2; m++)
		for (int d = 0; d < 3; d++)
		{
			free(st->bval[m][d]);
			free(st->bcond[m][d]);
		}
	free(st->ivox);
	free(st->ival);
	free(st->icond);
	for (int d = 0; d < 3; d++)
	{
		free(st->b[d]);
		free(st->c[d]);
		free(st->e[d]);
	}
	free(st->num);
	free(st->den);
	free(st->fac);
	free(st->rnum);
	free(st->rden);
	free(st->rfac);
	free(st->ballots);
	free(st);
}

void OR_FN(or_grid_shape)(const or_state* st, int* out3)
{
	for (int d = 0; d < 3; d++)
		out3[d] = st->grid[d];
}
or_real* OR_FN(or_densities)(or_state* st) { return st->dens; }
or_real* OR_FN(or_gradients)(or_state* st) { return st->grad; }

void OR_FN(or_set_boundary)(or_state* st, int dim, int is_max, const or_real* values, const unsigned char* conds)
{
	free(st->bval[is_max][dim]);
	free(st->bcond[is_max][dim]);
	st->bval[is_max][dim] = NULL;
	st->bcond[is_max][dim] = NULL;
	if (!values)
		return;
	st->bval[is_max][dim] = (or_real*)malloc(sizeof(or_real) * st->S);
	st->bcond[is_max][dim] = (unsigned char*)malloc(st->S);
	memcpy(st->bval[is_max][dim], values, sizeof(or_real) * st->S);
	memcpy(st->bcond[is_max][dim], conds, st->S);
}

void OR_FN(or_set_interior)(or_state* st, int count, const int* voxels, const or_real* values,
							const unsigned char* conds)
{
	free(st->ivox);
	free(st->ival);
	free(st->icond);
	st->ivox = NULL;
	st->ival = NULL;
	st->icond = NULL;
	st->n_interior = count;
	if (count <= 0)
		return;
	st->ivox = (int*)malloc(sizeof(int) * (size_t)count * st->dims);
	st->ival = (or_real*)malloc(sizeof(or_real) * (size_t)count * st->S);
	st->icond = (unsigned char*)malloc((size_t)count * st->S);
	memcpy(st->ivox, voxels, sizeof(int) * (size_t)count * st->dims);
	memcpy(st->ival, values, sizeof(or_real) * (size_t)count * st->S);
	memcpy(st->icond, conds, (size_t)count * st->S);
}

void OR_FN(or_set_agents)(or_state* st, int count, const or_real* secretion_rates, const or_real* saturation_densities,
						  const or_real* uptake_rates, const or_real* net_export_rates, or_real* internalized,
						  const or_real* fraction_released_at_death, const or_real* volumes, const or_real* positions)
{
	st->A = count;
	st->sec = secretion_rates;
	st->sat = saturation_densities;
	st->upt = uptake_rates;
	st->nex = net_export_rates;
	st->intern = internalized;
	st->frac = fraction_released_at_death;
	st->vol = volumes;
	st->pos = positions;
}

void OR_FN(or_set_compute_internalized)(or_state* st, int flag) { st->compute_internalized = flag; }

void OR_FN(or_set_threads)(int threads)
{
	if (threads > 0)
		omp_set_num_threads(threads);
}

/* ------------------------------------------------------------------------------------------ */
/* diffusion coefficients — host/diffusion_solver.cpp:14-44 (substrate factor), 72-142        */
/* ------------------------------------------------------------------------------------------ */

static int lcm_i(int a, int b)
{
	int high = a > b ? a : b;
	int low = a < b ? a : b;
	int ret = high;
	while (ret % low != 0)
		ret += high;
	return ret;
}

static int substrate_factor(int S)
{
	/* host/diffusion_solver.cpp:36-43: copies so that S*copies*bits is a multiple of a 512-bit register */
	int register_size = 512;
	int substrates_size = S * (int)sizeof(or_real) * 8;
	int multiple = lcm_i(substrates_size, register_size);
	while (multiple > 10 * register_size && multiple > substrates_size)
		multiple -= substrates_size;
	return multiple / substrates_size;
}

/* host/diffusion_solver.cpp:72-142 (layout [i][copies][s]) */
static void precompute_values(or_real** pb, or_real** pc, or_real** pe, int shape, int dims, int n,
							  const or_state* st, int copies)
{
	const int S = st->S;
	const or_real dt = st->dt;
	if (n == 1)
	{
		or_real* b = (or_real*)malloc(sizeof(or_real) * S * copies);
		for (int x = 0; x < copies; x++)
			for (int s = 0; s < S; s++)
				b[x * S + s] = 1 / (1 + st->decay[s] * dt / dims);
		*pb = b;
		*pc = NULL;
		*pe = NULL;
		return;
	}
	const int L = S * copies;
	or_real* b = (or_real*)malloc(sizeof(or_real) * (size_t)n * L);
	or_real* e = (or_real*)malloc(sizeof(or_real) * (size_t)(n - 1) * L);
	or_real* c = (or_real*)malloc(sizeof(or_real) * L);

	for (int x = 0; x < copies; x++)
		for (int s = 0; s < S; s++)
			c[x * S + s] = -dt * st->D[s] / (shape * shape);

	int ends[2] = { 0, n - 1 };
	for (int k = 0; k < 2; k++)
		for (int x = 0; x < copies; x++)
			for (int s = 0; s < S; s++)
				b[(size_t)ends[k] * L + x * S + s] =
					1 + st->decay[s] * dt / dims + dt * st->D[s] / (shape * shape);
	for (int i = 1; i < n - 1; i++)
		for (int x = 0; x < copies; x++)
			for (int s = 0; s < S; s++)
				b[(size_t)i * L + x * S + s] = 1 + st->decay[s] * dt / dims + 2 * dt * st->D[s] / (shape * shape);

	for (int x = 0; x < copies; x++)
		for (int s = 0; s < S; s++)
			b[x * S + s] = 1 / b[x * S + s];

	for (int i = 1; i < n; i++)
		for (int x = 0; x < copies; x++)
			for (int s = 0; s < S; s++)
			{
				const int k = x * S + s;
				b[(size_t)i * L + k] = 1 / (b[(size_t)i * L + k] - c[k] * c[k] * b[(size_t)(i - 1) * L + k]);
				e[(size_t)(i - 1) * L + k] = c[k] * b[(size_t)(i - 1) * L + k];
			}
	*pb = b;
	*pc = c;
	*pe = e;
}

int OR_FN(or_diffusion_coefficients)(const or_state* st, int axis, or_real* b_out, or_real* c_out, or_real* e_out)
{
	if (!st->initialized || axis < 0 || axis >= st->dims)
		return fail("or_diffusion_coefficients: not initialized or bad axis");
	const int n = st->grid[axis];
	const int L = st->S * st->copies[axis];
	/* copy out the first copy of each [i][copy][s] row */
	for (int i = 0; i < n; i++)
		for (int s = 0; s < st->S; s++)
		{
			if (b_out)
				b_out[(size_t)i * st->S + s] = st->b[axis][(size_t)i * L + s];
			if (e_out && i < n - 1)
				e_out[(size_t)i * st->S + s] = st->e[axis][(size_t)i * L + s];
		}
	if (c_out && n > 1)
		for (int s = 0; s < st->S; s++)
			c_out[s] = st->c[axis][s];
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* cell solver resize/initialize — host/cell_solver.cpp:348-373                               */
/* ------------------------------------------------------------------------------------------ */

static void cell_resize(or_state* st)
{
	size_t need = (size_t)st->A * st->S;
	if (need > st->cap || !st->num)
	{
		size_t cap = need > 0 ? need : 1;
		free(st->num);
		free(st->den);
		free(st->fac);
		free(st->rnum);
		free(st->rden);
		free(st->rfac);
		st->num = (or_real*)calloc(cap, sizeof(or_real));
		st->den = (or_real*)calloc(cap, sizeof(or_real));
		st->fac = (or_real*)calloc(cap, sizeof(or_real));
		st->rnum = (or_real*)calloc(cap, sizeof(or_real));
		st->rden = (or_real*)calloc(cap, sizeof(or_real));
		st->rfac = (or_real*)calloc(cap, sizeof(or_real));
		st->cap = cap;
	}
}

int OR_FN(or_initialize)(or_state* st)
{
	/* cell.initialize: host/cell_solver.cpp:366-373 */
	cell_resize(st);
	free(st->ballots);
	/* std::make_unique<std::atomic<index_t>[]>(V) value-initialises the ballots to 0 */
	st->ballots = (int*)calloc(voxel_count(st), sizeof(int));
	st->is_conflict = 0;

	/* diffusion.initialize: host/diffusion_solver.cpp:29-58 */
	st->substrate_factor = substrate_factor(st->S);
	for (int d = 0; d < 3; d++)
	{
		free(st->b[d]);
		free(st->c[d]);
		free(st->e[d]);
		st->b[d] = st->c[d] = st->e[d] = NULL;
	}
	for (int d = 0; d < st->dims; d++)
	{
		st->copies[d] = d == 0 ? 1 : st->substrate_factor;
		precompute_values(&st->b[d], &st->c[d], &st->e[d], st->vs[d], st->dims, st->grid[d], st, st->copies[d]);
	}
	st->initialized = 1;
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Dirichlet — host/dirichlet_solver.cpp:22-144 (orphaned worksharing)                         */
/* ------------------------------------------------------------------------------------------ */

/* host/dirichlet_solver.cpp:46-59: one face, every substrate whose condition is set */
static void solve_boundary(or_state* st, int dim, int is_max)
{
	const or_real* vals = st->bval[is_max][dim];
	const unsigned char* conds = st->bcond[is_max][dim];
	if (!vals)
		return;
	const int X = st->grid[0], Y = st->grid[1], Z = st->grid[2], S = st->S;
	const int fixed = is_max ? st->grid[dim] - 1 : 0;
	if (dim == 0)
	{
#pragma omp for
		for (int z = 0; z < Z; z++)
			for (int y = 0; y < Y; y++)
			{
				or_real* p = st->dens + (((size_t)z * Y + y) * X + fixed) * S;
				for (int s = 0; s < S; s++)
					if (conds[s])
						p[s] = vals[s];
			}
	}
	else if (dim == 1)
	{
#pragma omp for
		for (int z = 0; z < Z; z++)
			for (int x = 0; x < X; x++)
			{
				or_real* p = st->dens + (((size_t)z * Y + fixed) * X + x) * S;
				for (int s = 0; s < S; s++)
					if (conds[s])
						p[s] = vals[s];
			}
	}
	else
	{
#pragma omp for
		for (int y = 0; y < Y; y++)
			for (int x = 0; x < X; x++)
			{
				or_real* p = st->dens + (((size_t)fixed * Y + y) * X + x) * S;
				for (int s = 0; s < S; s++)
					if (conds[s])
						p[s] = vals[s];
			}
	}
}

/* host/dirichlet_solver.cpp:22-44 */
static void solve_interior(or_state* st)
{
	if (st->n_interior == 0)
		return;
	const int X = st->grid[0], Y = st->grid[1], S = st->S, dims = st->dims;
#pragma omp for
	for (int i = 0; i < st->n_interior; i++)
	{
		const int* v = st->ivox + (size_t)i * dims;
		int x = v[0], y = dims > 1 ? v[1] : 0, z = dims > 2 ? v[2] : 0;
		or_real* p = st->dens + (((size_t)z * Y + y) * X + x) * S;
		for (int s = 0; s < S; s++)
			if (st->icond[(size_t)i * S + s])
				p[s] = st->ival[(size_t)i * S + s];
	}
}

/* host/dirichlet_solver.cpp:61-110 then :122-144 — faces x-,x+,y-,y+,z-,z+ then the interior list */
static void dirichlet_orphan(or_state* st)
{
	for (int d = 0; d < st->dims; d++)
	{
		solve_boundary(st, d, 0);
		solve_boundary(st, d, 1);
	}
	solve_interior(st);
}

int OR_FN(or_dirichlet_solve)(or_state* st)
{
#pragma omp parallel
	dirichlet_orphan(st);
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Thomas sweeps — host/diffusion_solver.cpp:297-480                                           */
/* ------------------------------------------------------------------------------------------ */

/* host/diffusion_solver.cpp:297-334: one x line of n voxels × S substrates, stride S */
static void solve_slice_x(or_real* restrict d, const or_real* restrict b, const or_real* restrict c,
						  const or_real* restrict e, int n, int S)
{
	for (int i = 1; i < n; i++)
		for (int s = 0; s < S; s++)
			d[(size_t)i * S + s] = d[(size_t)i * S + s] - e[(size_t)(i - 1) * S + s] * d[(size_t)(i - 1) * S + s];
	for (int s = 0; s < S; s++)
		d[(size_t)(n - 1) * S + s] = d[(size_t)(n - 1) * S + s] * b[(size_t)(n - 1) * S + s];
	for (int i = n - 2; i >= 0; i--)
		for (int s = 0; s < S; s++)
			d[(size_t)i * S + s] = (d[(size_t)i * S + s] - c[s] * d[(size_t)(i + 1) * S + s]) * b[(size_t)i * S + s];
}

/* host/diffusion_solver.cpp:336-376: 1-D x sweep parallel over substrates */
static void solve_slice_x_omp(or_real* restrict d, const or_real* restrict b, const or_real* restrict c,
							  const or_real* restrict e, int n, int S)
{
#pragma omp for
	for (int s = 0; s < S; s++)
		for (int i = 1; i < n; i++)
			d[(size_t)i * S + s] = d[(size_t)i * S + s] - e[(size_t)(i - 1) * S + s] * d[(size_t)(i - 1) * S + s];
#pragma omp for
	for (int s = 0; s < S; s++)
		d[(size_t)(n - 1) * S + s] = d[(size_t)(n - 1) * S + s] * b[(size_t)(n - 1) * S + s];
#pragma omp for
	for (int s = 0; s < S; s++)
		for (int i = n - 2; i >= 0; i--)
			d[(size_t)i * S + s] = (d[(size_t)i * S + s] - c[s] * d[(size_t)(i + 1) * S + s]) * b[(size_t)i * S + s];
}

/*
 * host/diffusion_solver.cpp:378-414: sweep along an axis whose lines are `row` contiguous elements apart,
 * vectorised over the `row` = X*S elements of a plane row.  Coefficients are laid out [i][L] with
 * L = S*copies, so element k of the row uses coefficient k % L.
 */
static void solve_slice_plane(or_real* restrict d, const or_real* restrict b, const or_real* restrict c,
							  const or_real* restrict e, int n, size_t row, size_t stride, int L)
{
	for (int i = 1; i < n; i++)
	{
		or_real* cur = d + (size_t)i * stride;
		const or_real* prev = d + (size_t)(i - 1) * stride;
		const or_real* ee = e + (size_t)(i - 1) * L;
		for (size_t k0 = 0; k0 < row; k0 += L)
		{
			size_t len = row - k0 < (size_t)L ? row - k0 : (size_t)L;
			for (size_t k = 0; k < len; k++)
				cur[k0 + k] = cur[k0 + k] - ee[k] * prev[k0 + k];
		}
	}
	{
		or_real* last = d + (size_t)(n - 1) * stride;
		const or_real* bb = b + (size_t)(n - 1) * L;
		for (size_t k0 = 0; k0 < row; k0 += L)
		{
			size_t len = row - k0 < (size_t)L ? row - k0 : (size_t)L;
			for (size_t k = 0; k < len; k++)
				last[k0 + k] = last[k0 + k] * bb[k];
		}
	}
	for (int i = n - 2; i >= 0; i--)
	{
		or_real* cur = d + (size_t)i * stride;
		const or_real* next = d + (size_t)(i + 1) * stride;
		const or_real* bb = b + (size_t)i * L;
		for (size_t k0 = 0; k0 < row; k0 += L)
		{
			size_t len = row - k0 < (size_t)L ? row - k0 : (size_t)L;
			for (size_t k = 0; k < len; k++)
				cur[k0 + k] = (cur[k0 + k] - c[k] * next[k0 + k]) * bb[k];
		}
	}
}

/*
 * host/diffusion_solver.cpp:416-480: sweep along the outermost axis of `planes` rows of `row` elements,
 * every step i parallel over rows (omp for nowait), then a barrier.
 */
static void solve_slice_outer_omp(or_real* restrict d, const or_real* restrict b, const or_real* restrict c,
								  const or_real* restrict e, int n, int rows, size_t row, int L)
{
	const size_t stride = (size_t)rows * row;
	for (int i = 1; i < n; i++)
	{
		const or_real* ee = e + (size_t)(i - 1) * L;
#pragma omp for nowait
		for (int j = 0; j < rows; j++)
		{
			or_real* cur = d + (size_t)i * stride + (size_t)j * row;
			const or_real* prev = cur - stride;
			for (size_t k0 = 0; k0 < row; k0 += L)
			{
				size_t len = row - k0 < (size_t)L ? row - k0 : (size_t)L;
				for (size_t k = 0; k < len; k++)
					cur[k0 + k] = cur[k0 + k] - ee[k] * prev[k0 + k];
			}
		}
	}
#pragma omp barrier
	{
		const or_real* bb = b + (size_t)(n - 1) * L;
#pragma omp for
		for (int j = 0; j < rows; j++)
		{
			or_real* cur = d + (size_t)(n - 1) * stride + (size_t)j * row;
			for (size_t k0 = 0; k0 < row; k0 += L)
			{
				size_t len = row - k0 < (size_t)L ? row - k0 : (size_t)L;
				for (size_t k = 0; k < len; k++)
					cur[k0 + k] = cur[k0 + k] * bb[k];
			}
		}
	}
	for (int i = n - 2; i >= 0; i--)
	{
		const or_real* bb = b + (size_t)i * L;
#pragma omp for nowait
		for (int j = 0; j < rows; j++)
		{
			or_real* cur = d + (size_t)i * stride + (size_t)j * row;
			const or_real* next = cur + stride;
			for (size_t k0 = 0; k0 < row; k0 += L)
			{
				size_t len = row - k0 < (size_t)L ? row - k0 : (size_t)L;
				for (size_t k = 0; k < len; k++)
					cur[k0 + k] = (cur[k0 + k] - c[k] * next[k0 + k]) * bb[k];
			}
		}
	}
#pragma omp barrier
}

/* n == 1 along an axis: the reference's sweep reduces to d *= b (host/diffusion_solver.cpp:76-85, 317-322) */
static void scale_all_omp(or_state* st, const or_real* b)
{
	const size_t V = voxel_count(st);
	const int S = st->S;
#pragma omp for
	for (long long v = 0; v < (long long)V; v++)
		for (int s = 0; s < S; s++)
			st->dens[(size_t)v * S + s] = st->dens[(size_t)v * S + s] * b[s];
}

/* host/diffusion_solver.cpp:482-544 */
static void diffusion_orphan(or_state* st)
{
	const int X = st->grid[0], Y = st->grid[1], Z = st->grid[2], S = st->S;
	const size_t row = (size_t)X * S;

	dirichlet_orphan(st);

	/* x sweep */
	if (X == 1)
		scale_all_omp(st, st->b[0]);
	else if (st->dims == 1)
		solve_slice_x_omp(st->dens, st->b[0], st->c[0], st->e[0], X, S);
	else
	{
#pragma omp for
		for (int zy = 0; zy < Z * Y; zy++)
			solve_slice_x(st->dens + (size_t)zy * row, st->b[0], st->c[0], st->e[0], X, S);
	}
	dirichlet_orphan(st);
	if (st->dims == 1)
		return;

	/* y sweep */
	const int Ly = S * st->copies[1];
	if (Y == 1)
		scale_all_omp(st, st->b[1]);
	else if (st->dims == 2)
		solve_slice_outer_omp(st->dens, st->b[1], st->c[1], st->e[1], Y, 1, row, Ly);
	else
	{
#pragma omp for
		for (int z = 0; z < Z; z++)
			solve_slice_plane(st->dens + (size_t)z * Y * row, st->b[1], st->c[1], st->e[1], Y, row, row, Ly);
	}
	dirichlet_orphan(st);
	if (st->dims == 2)
		return;

	/* z sweep */
	const int Lz = S * st->copies[2];
	if (Z == 1)
		scale_all_omp(st, st->b[2]);
	else
		solve_slice_outer_omp(st->dens, st->b[2], st->c[2], st->e[2], Z, Y, row, Lz);
	dirichlet_orphan(st);
}

int OR_FN(or_diffusion_solve)(or_state* st)
{
	if (!st->initialized)
		return fail("or_diffusion_solve: not initialized");
#pragma omp parallel
	diffusion_orphan(st);
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* gradient — host/gradient_solver.cpp:10-291                                                 */
/* ------------------------------------------------------------------------------------------ */

/*
 * Central difference along one axis; the reference multiplies by float reciprocals
 * (host/gradient_solver.cpp:17-18: 1.f / dx and 1.f / (2 dx)) even in double builds.
 * Gradient element (d, voxel v, s) is stored at gidx(d, v, s).
 */
static inline size_t gidx(const or_state* st, int d, size_t v, int s)
{
	if (st->dims == 3)
		return ((size_t)d * voxel_count(st) + v) * st->S + s;
	return (v * st->S + s) * st->dims + d;
}

static void gradient_axis(or_state* st, int d)
{
	const int X = st->grid[0], Y = st->grid[1], Z = st->grid[2], S = st->S;
	const int n = st->grid[d];
	const or_real border_factor = 1.f / st->vs[d];
	const or_real center_factor = 1.f / (st->vs[d] * 2);
	const size_t astride = d == 0 ? 1 : d == 1 ? (size_t)X : (size_t)X * Y; /* voxel stride along the axis */
	/* distribute over z for x and y (host/gradient_solver.cpp:232-247), over the z index for z (:185-222) */
#pragma omp for
	for (int z = 0; z < Z; z++)
		for (int y = 0; y < Y; y++)
			for (int x = 0; x < X; x++)
			{
				const int i = d == 0 ? x : d == 1 ? y : z;
				const size_t v = ((size_t)z * Y + y) * X + x;
				for (int s = 0; s < S; s++)
				{
					or_real val;
					if (i == 0)
						val = (st->dens[(v + astride) * S + s] - st->dens[v * S + s]) * border_factor;
					else if (i == n - 1)
						val = (st->dens[v * S + s] - st->dens[(v - astride) * S + s]) * border_factor;
					else
						val = (st->dens[(v + astride) * S + s] - st->dens[(v - astride) * S + s]) * center_factor;
					st->grad[gidx(st, d, v, s)] = val;
				}
			}
}

static int gradient_check(const or_state* st)
{
	for (int d = 0; d < st->dims; d++)
		if (st->grid[d] < 2)
			return fail("gradient: every axis needs at least 2 voxels");
	return 0;
}

static void gradient_orphan(or_state* st)
{
	for (int d = 0; d < st->dims; d++)
		gradient_axis(st, d);
}

int OR_FN(or_gradient_solve)(or_state* st)
{
	if (gradient_check(st))
		return -1;
#pragma omp parallel
	gradient_orphan(st);
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* bulk supply / uptake — host/bulk_solver.cpp:23-35, 98-120                                  */
/* ------------------------------------------------------------------------------------------ */

/* The reference asks per-voxel callbacks for S, U and T at every solve; the oracle takes the values they
 * return as [V][S] arrays (the densities' layout) and applies, per voxel and substrate,
 *   D = (D + dt*S*T) / (1 + dt*(U + S))
 * with the reference's operation order: ((dt * S) * T) in the numerator. */
int OR_FN(or_bulk_solve)(or_state* st, const or_real* supply, const or_real* uptake, const or_real* target)
{
	const long long n = (long long)st->grid[0] * st->grid[1] * st->grid[2] * st->S;
	const or_real dt = st->dt;
#pragma omp parallel for schedule(static)
	for (long long i = 0; i < n; i++)
		st->dens[i] = (st->dens[i] + dt * supply[i] * target[i]) / (1 + dt * (uptake[i] + supply[i]));
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* secretion / uptake — host/cell_solver.cpp:15-346                                           */
/* ------------------------------------------------------------------------------------------ */

/* mesh.cpp:47-53: truncate (p - min) / dx toward zero; returns the linear voxel or -1 when outside. */
static inline long long agent_voxel(const or_state* st, int i)
{
	int idx[3] = { 0, 0, 0 };
	for (int d = 0; d < st->dims; d++)
	{
		or_real q = (st->pos[(size_t)i * st->dims + d] - st->mins[d]) / st->vs[d];
		if (!(q > -1) || !(q < st->grid[d]))
			return -1;
		idx[d] = (int)q;
	}
	return ((long long)idx[2] * st->grid[1] + idx[1]) * st->grid[0] + idx[0];
}

void OR_FN(or_voxel_positions)(const or_state* st, int count, const or_real* positions, int* out3)
{
	for (int i = 0; i < count; i++)
		for (int d = 0; d < 3; d++)
		{
			if (d < st->dims)
			{
				or_real q = (positions[(size_t)i * st->dims + d] - st->mins[d]) / st->vs[d];
				out3[i * 3 + d] = (q > -1 && q < st->grid[d]) ? (int)q : -1;
			}
			else
				out3[i * 3 + d] = 0;
		}
}

/* host/cell_solver.cpp:63-85 */
static void compute_intermediates(or_state* st)
{
	const int S = st->S;
	const or_real voxel_volume = (or_real)(st->vs[0] * st->vs[1] * st->vs[2]);
	const or_real time_step = st->dt;
#pragma omp for
	for (int i = 0; i < st->A; i++)
		for (int s = 0; s < S; s++)
		{
			const size_t k = (size_t)i * S + s;
			st->num[k] = st->sec[k] * st->sat[k] * time_step * st->vol[i] / voxel_volume;
			st->den[k] = (st->upt[k] + st->sec[k]) * time_step * st->vol[i] / voxel_volume;
			st->fac[k] = st->nex[k] * time_step / voxel_volume;
		}
}

/* host/cell_solver.cpp:35-61 */
static void clear_ballots(or_state* st)
{
	const int S = st->S;
#pragma omp for
	for (int i = 0; i < st->A; i++)
	{
		long long v = agent_voxel(st, i);
		if (v >= 0)
			__atomic_store_n(&st->ballots[v], -1, __ATOMIC_RELAXED);
		for (int s = 0; s < S; s++)
		{
			st->rnum[(size_t)i * S + s] = 0;
			st->rden[(size_t)i * S + s] = 0;
			st->rfac[(size_t)i * S + s] = 0;
		}
	}
}

/* host/cell_solver.cpp:87-135 */
static void ballot_and_sum(or_state* st)
{
	const int S = st->S;
#pragma omp for
	for (int i = 0; i < st->A; i++)
	{
		long long v = agent_voxel(st, i);
		if (v < 0)
			continue;
		int expected = -1;
		int success = __atomic_compare_exchange_n(&st->ballots[v], &expected, i, 0, __ATOMIC_ACQ_REL,
												  __ATOMIC_ACQUIRE);
		if (success)
		{
			for (int s = 0; s < S; s++)
			{
				const size_t k = (size_t)i * S + s;
				const or_real dplus = st->den[k] + 1;
#pragma omp atomic
				st->rnum[k] += st->num[k];
#pragma omp atomic
				st->rden[k] += dplus;
#pragma omp atomic
				st->rfac[k] += st->fac[k];
			}
		}
		else
		{
			__atomic_store_n(&st->is_conflict, 1, __ATOMIC_RELAXED);
			for (int s = 0; s < S; s++)
			{
				const size_t k = (size_t)i * S + s;
				const size_t w = (size_t)expected * S + s;
#pragma omp atomic
				st->rnum[w] += st->num[k];
#pragma omp atomic
				st->rden[w] += st->den[k];
#pragma omp atomic
				st->rfac[w] += st->fac[k];
			}
		}
	}
}

/* host/cell_solver.cpp:197-245 (with compute_internalized :137-152, compute_densities :154-171,
 * compute_fused :173-195) */
static void compute_result(or_state* st, int with_internalized, int is_conflict)
{
	const int S = st->S;
	const int voxel_volume = st->vs[0] * st->vs[1] * st->vs[2];
	if (with_internalized && !is_conflict)
	{
#pragma omp for
		for (int i = 0; i < st->A; i++)
		{
			long long v = agent_voxel(st, i);
			if (v < 0)
				continue;
			or_real* D = st->dens + (size_t)v * S;
			or_real* I = st->intern + (size_t)i * S;
			for (int s = 0; s < S; s++)
			{
				const size_t k = (size_t)i * S + s;
				I[s] = I[s] - voxel_volume * (((1 - st->rden[k]) * D[s] + st->rnum[k]) / (st->rden[k]) + st->rfac[k]);
				D[s] = (D[s] + st->rnum[k]) / st->rden[k] + st->rfac[k];
			}
		}
		return;
	}
	if (with_internalized)
	{
		const or_real vv = (or_real)voxel_volume;
#pragma omp for
		for (int i = 0; i < st->A; i++)
		{
			long long v = agent_voxel(st, i);
			if (v < 0)
				continue;
			const or_real* D = st->dens + (size_t)v * S;
			or_real* I = st->intern + (size_t)i * S;
			for (int s = 0; s < S; s++)
			{
				const size_t k = (size_t)i * S + s;
				I[s] = I[s] - vv * (((-st->den[k]) * D[s] + st->num[k]) / (1 + st->den[k]) + st->fac[k]);
			}
		}
	}
#pragma omp for
	for (int i = 0; i < st->A; i++)
	{
		long long v = agent_voxel(st, i);
		if (v < 0)
			continue;
		int ballot = __atomic_load_n(&st->ballots[v], __ATOMIC_RELAXED);
		if (ballot != i)
			continue;
		or_real* D = st->dens + (size_t)v * S;
		for (int s = 0; s < S; s++)
		{
			const size_t k = (size_t)i * S + s;
			D[s] = (D[s] + st->rnum[k]) / st->rden[k] + st->rfac[k];
		}
	}
}

/* host/cell_solver.cpp:247-299 */
static void cell_orphan(or_state* st, int recompute)
{
#pragma omp single
	if (recompute)
	{
		cell_resize(st);
		st->is_conflict = 0;
	}
	if (recompute)
	{
		compute_intermediates(st);
		clear_ballots(st);
		ballot_and_sum(st);
	}
	compute_result(st, st->compute_internalized, __atomic_load_n(&st->is_conflict, __ATOMIC_RELAXED));
}

int OR_FN(or_cell_simulate)(or_state* st, int recompute)
{
	if (!st->initialized)
		return fail("or_cell_simulate: not initialized");
	if (st->A > 0 && (!st->pos || !st->vol))
		return fail("or_cell_simulate: agents not set");
#pragma omp parallel
	cell_orphan(st, recompute);
	return 0;
}

/* host/cell_solver.cpp:301-346 */
int OR_FN(or_cell_release)(or_state* st, int index)
{
	if (!st->compute_internalized)
		return 0;
	if (index < 0 || index >= st->A)
		return fail("or_cell_release: index out of range");
	long long v = agent_voxel(st, index);
	if (v < 0)
		return 0;
	const or_real voxel_volume = (or_real)(st->vs[0] * st->vs[1] * st->vs[2]);
	or_real* D = st->dens + (size_t)v * st->S;
	or_real* I = st->intern + (size_t)index * st->S;
	for (int s = 0; s < st->S; s++)
	{
		D[s] = D[s] + I[s] * st->frac[(size_t)index * st->S + s] / voxel_volume;
		I[s] = 0;
	}
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* one timestep, one parallel region (main.cpp:64-107), order secretion -> diffusion -> gradient */
/* ------------------------------------------------------------------------------------------ */

int OR_FN(or_timestep)(or_state* st, int recompute, int with_gradient, double* phase_ms)
{
	if (!st->initialized)
		return fail("or_timestep: not initialized");
	if (with_gradient && gradient_check(st))
		return -1;
	double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
#pragma omp parallel
	{
#pragma omp master
		t0 = omp_get_wtime();
		cell_orphan(st, recompute);
#pragma omp barrier
#pragma omp master
		t1 = omp_get_wtime();
		diffusion_orphan(st);
#pragma omp barrier
#pragma omp master
		t2 = omp_get_wtime();
		if (with_gradient)
			gradient_orphan(st);
#pragma omp barrier
#pragma omp master
		t3 = omp_get_wtime();
	}
	if (phase_ms)
	{
		phase_ms[0] = (t1 - t0) * 1e3;
		phase_ms[1] = (t2 - t1) * 1e3;
		phase_ms[2] = (t3 - t2) * 1e3;
	}
	return 0;
}
