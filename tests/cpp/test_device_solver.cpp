// The reference's device test-suite cases (test/test_device.cpp, expectations from test/test.cpp) run through
// the C++ drop-in facade (biofvm::solver = solvers::device::solver over libparabiofvm_hip.so).
//
// Comparison: the reference uses EXPECT_FLOAT_EQ (4 float ULPs) against its host float path; the device
// path here solves the same systems with a partitioned Thomas algorithm and FP64-derived coefficients, so
// a value passes when it is within 4 ULPs OR within 1e-5 relative of the expectation (the parity bar of
// the FP32 device path).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "agent_container.h"
#include "microenvironment.h"
#include "microenvironment_builder.h"
#include "solver.h"

using namespace biofvm;

static int g_failures = 0, g_checks = 0;
static std::string g_test;

static int64_t ulp_distance(float a, float b)
{
	int32_t ia, ib;
	std::memcpy(&ia, &a, 4);
	std::memcpy(&ib, &b, 4);
	int64_t la = ia < 0 ? (int64_t)INT32_MIN - ia : ia;
	int64_t lb = ib < 0 ? (int64_t)INT32_MIN - ib : ib;
	return la > lb ? la - lb : lb - la;
}

static void expect_eq(double got, double exp, const char* file, int line)
{
	g_checks++;
	const bool ok = ulp_distance((float)got, (float)exp) <= 4 || std::fabs(got - exp) <= 1e-5 * std::fabs(exp);
	if (!ok)
	{
		if (g_failures < 50)
			std::printf("  FAIL %s (%s:%d): got %.9g expected %.9g\n", g_test.c_str(), file, line, got, exp);
		g_failures++;
	}
}
#define EXPECT_FLOAT_EQ(a, b) expect_eq((a), (b), __FILE__, __LINE__)

// A running difference (an agent's internalized amount) is compared relative to the sum of the magnitudes of
// the terms it accumulated, the componentwise error bound of a floating-point sum.
static void expect_near_scaled(double got, double exp, double scale, const char* file, int line)
{
	g_checks++;
	if (!(std::fabs(got - exp) <= 1e-5 * (std::fabs(exp) + scale)))
	{
		if (g_failures < 50)
			std::printf("  FAIL %s (%s:%d): got %.9g expected %.9g (term scale %.3g)\n", g_test.c_str(), file, line,
						got, exp, scale);
		g_failures++;
	}
}

// ---- fixtures (test/utils.cpp:10-250 of the reference, restated) ----
static microenvironment default_microenv(cartesian_mesh mesh)
{
	real_t init[2] = { 1, 1 };
	microenvironment m(mesh, 2, 5, init);
	m.diffusion_coefficients = std::make_unique<real_t[]>(2);
	m.diffusion_coefficients[0] = 4;
	m.diffusion_coefficients[1] = 2;
	m.decay_rates = std::make_unique<real_t[]>(2);
	m.decay_rates[0] = 5;
	m.decay_rates[1] = 3;
	return m;
}

static microenvironment biorobots_microenv(cartesian_mesh mesh)
{
	real_t init[2] = { 0, 0 };
	microenvironment m(mesh, 2, 0.01f, init);
	m.diffusion_coefficients = std::make_unique<real_t[]>(2);
	m.diffusion_coefficients[0] = 1000;
	m.diffusion_coefficients[1] = 1000;
	m.decay_rates = std::make_unique<real_t[]>(2);
	m.decay_rates[0] = 0.1f;
	m.decay_rates[1] = 0.4f;
	return m;
}

// interior Dirichlet on substrate 0 only
static void add_dirichlet_at(microenvironment& m, const std::vector<point_t<index_t, 3>>& at,
							 const std::vector<real_t>& values)
{
	const index_t S = m.substrates_count;
	m.dirichlet_interior_voxels_count = (index_t)at.size();
	m.dirichlet_interior_voxels = std::make_unique<index_t[]>(at.size() * m.mesh.dims);
	m.dirichlet_interior_values = std::make_unique<real_t[]>(at.size() * S);
	m.dirichlet_interior_conditions = std::make_unique<bool[]>(at.size() * S);
	for (std::size_t i = 0; i < at.size(); i++)
	{
		for (index_t d = 0; d < m.mesh.dims; d++)
			m.dirichlet_interior_voxels[i * m.mesh.dims + d] = at[i][d];
		for (index_t s = 0; s < S; s++)
		{
			m.dirichlet_interior_values[i * S + s] = s == 0 ? values[i] : 0;
			m.dirichlet_interior_conditions[i * S + s] = s == 0;
		}
	}
}

// boundary Dirichlet on substrate 0 of one face; other substrates 42 with no condition
static void add_boundary_dirichlet(microenvironment& m, index_t dim, bool min, real_t value)
{
	auto& vals = min ? m.dirichlet_min_boundary_values[dim] : m.dirichlet_max_boundary_values[dim];
	auto& conds = min ? m.dirichlet_min_boundary_conditions[dim] : m.dirichlet_max_boundary_conditions[dim];
	if (!vals)
	{
		vals = std::make_unique<real_t[]>(m.substrates_count);
		conds = std::make_unique<bool[]>(m.substrates_count);
		for (index_t s = 0; s < m.substrates_count; s++)
		{
			vals[s] = 42;
			conds[s] = false;
		}
	}
	vals[0] = value;
	conds[0] = true;
}

static void set_default_agent_values(agent* a, index_t offset, index_t volume, point_t<real_t, 3> pos, index_t dims)
{
	a->secretion_rates()[0] = offset + 100;
	a->secretion_rates()[1] = 0;
	a->uptake_rates()[0] = offset + 200;
	a->uptake_rates()[1] = 0;
	a->saturation_densities()[0] = offset + 300;
	a->saturation_densities()[1] = 0;
	a->net_export_rates()[0] = offset + 400;
	a->net_export_rates()[1] = 0;
	a->volume() = volume;
	for (index_t i = 0; i < dims; i++)
		a->position()[i] = pos[i];
}

static real_t dens(microenvironment& m, index_t x, index_t y, index_t z, index_t s)
{
	const auto& g = m.mesh.grid_shape;
	return m.substrate_densities[(((std::size_t)z * g[1] + y) * g[0] + x) * m.substrates_count + s];
}

static solver& the_solver()
{
	static solver s;
	return s;
}

// test/test_device.cpp:25-32: store, run, load
static void runit(solver& s, microenvironment& m, const std::function<void()>& f)
{
	s.store_data_to_solver(m);
	f();
	s.load_data_from_solver(m);
}

// ---- cases ----
static void diffusion_uniform(index_t dims, point_t<index_t, 3> maxs, point_t<index_t, 3> vs, double e0, double e1)
{
	cartesian_mesh mesh(dims, { 0, 0, 0 }, maxs, vs);
	auto m = default_microenv(mesh);
	auto& s = the_solver();
	s.initialize(m);
	runit(s, m, [&] { s.diffusion.solve(m); });
	for (std::size_t v = 0; v < mesh.voxel_count(); v++)
	{
		EXPECT_FLOAT_EQ(m.substrate_densities[v * 2 + 0], e0);
		EXPECT_FLOAT_EQ(m.substrate_densities[v * 2 + 1], e1);
	}
}

static void diffusion_random(index_t dims, point_t<index_t, 3> maxs, const std::vector<double>& expected)
{
	cartesian_mesh mesh(dims, { 0, 0, 0 }, maxs, { 20, 20, 20 });
	auto m = biorobots_microenv(mesh);
	for (std::size_t i = 0; i < mesh.voxel_count() * 2; i++)
		m.substrate_densities[i] = (real_t)i;
	auto& s = the_solver();
	s.initialize(m);
	runit(s, m, [&] { s.diffusion.solve(m); });
	for (std::size_t i = 0; i < expected.size(); i++)
		EXPECT_FLOAT_EQ(m.substrate_densities[i], expected[i]);
}

static void dirichlet_one_cond_D3()
{
	cartesian_mesh mesh(3, { 0, 0, 0 }, { 60, 60, 60 }, { 20, 20, 20 });
	auto m = default_microenv(mesh);
	add_dirichlet_at(m, { { 1, 1, 1 } }, { 1000 });
	auto& s = the_solver();
	s.initialize(m);
	runit(s, m, [&] { s.diffusion.solve(m); });
	for (index_t z = 0; z < 3; z++)
		for (index_t y = 0; y < 3; y++)
			for (index_t x = 0; x < 3; x++)
				EXPECT_FLOAT_EQ(dens(m, x, y, z, 1), 0.0046296306);
	const double ring[3][3] = { { 0.0012301364, 0.001549035, 0.0012301364 },
								{ 0.0012637526, 0.566124, 0.0012637526 },
								{ 0.0012301364, 0.001549035, 0.0012301364 } };
	for (index_t z : { 0, 2 })
		for (index_t y = 0; y < 3; y++)
			for (index_t x = 0; x < 3; x++)
				EXPECT_FLOAT_EQ(dens(m, x, y, z, 0), ring[y][x]);
	const double mid[3][3] = { { 0.0012637526, 0.061110392, 0.0012637526 },
							   { 0.0075723953, 1000, 0.0075723953 },
							   { 0.0012637526, 0.061110392, 0.0012637526 } };
	for (index_t y = 0; y < 3; y++)
		for (index_t x = 0; x < 3; x++)
			EXPECT_FLOAT_EQ(dens(m, x, y, 1, 0), mid[y][x]);
}

static void dirichlet_multiple_cond_D2()
{
	cartesian_mesh mesh(2, { 0, 0, 0 }, { 60, 60, 0 }, { 20, 20, 0 });
	auto m = default_microenv(mesh);
	add_dirichlet_at(m, { { 0, 0, 0 }, { 1, 0, 0 }, { 2, 0, 0 } }, { 10, 10, 10 });
	auto& s = the_solver();
	s.initialize(m);
	runit(s, m, [&] { s.diffusion.solve(m); });
	for (index_t y = 0; y < 3; y++)
		for (index_t x = 0; x < 3; x++)
		{
			EXPECT_FLOAT_EQ(dens(m, x, y, 0, 1), 0.013840831);
			EXPECT_FLOAT_EQ(dens(m, x, y, 0, 0), y == 0 ? 10 : (y == 1 ? 0.0081802057 : 0.0054969075));
		}
}

static void gradient_D3()
{
	cartesian_mesh mesh(3, { 0, 0, 0 }, { 100, 100, 100 }, { 20, 20, 20 });
	auto m = default_microenv(mesh);
	auto f = [](index_t x, index_t y, index_t z, index_t s) { return (real_t)((x + y + z) * (x + y + z) + s); };
	for (index_t z = 0; z < 5; z++)
		for (index_t y = 0; y < 5; y++)
			for (index_t x = 0; x < 5; x++)
				for (index_t s = 0; s < 2; s++)
					m.substrate_densities[(((std::size_t)z * 5 + y) * 5 + x) * 2 + s] = f(x, y, z, s);
	auto& s = the_solver();
	s.initialize(m);
	runit(s, m, [&] { s.gradient.solve(m); });
	const std::size_t VS = 125 * 2;
	for (index_t z = 0; z < 5; z++)
		for (index_t y = 0; y < 5; y++)
			for (index_t x = 0; x < 5; x++)
				for (index_t sb = 0; sb < 2; sb++)
				{
					const index_t c[3] = { x, y, z };
					for (int d = 0; d < 3; d++)
					{
						index_t up[3] = { x, y, z }, dn[3] = { x, y, z };
						up[d] = c[d] + (c[d] == 4 ? 0 : 1);
						dn[d] = c[d] - (c[d] == 0 ? 0 : 1);
						const double div = (c[d] == 0 || c[d] == 4) ? 20. : 40.;
						const std::size_t e = (((std::size_t)z * 5 + y) * 5 + x) * 2 + sb;
						EXPECT_FLOAT_EQ(m.gradients[d * VS + e],
										(f(up[0], up[1], up[2], sb) - f(dn[0], dn[1], dn[2], sb)) / div);
					}
				}
}

static void boundaries_D3()
{
	cartesian_mesh mesh(3, { 0, 0, 0 }, { 100, 100, 100 }, { 20, 20, 20 });
	auto m = default_microenv(mesh);
	add_boundary_dirichlet(m, 0, true, 4);
	add_boundary_dirichlet(m, 0, false, 5);
	add_boundary_dirichlet(m, 1, true, 6);
	add_boundary_dirichlet(m, 1, false, 7);
	add_boundary_dirichlet(m, 2, true, 8);
	add_boundary_dirichlet(m, 2, false, 9);
	auto& s = the_solver();
	s.initialize(m);
	runit(s, m, [&] { s.diffusion.dirichlet.solve(m); });
	for (index_t z = 0; z < 5; z++)
		for (index_t y = 0; y < 5; y++)
			for (index_t x = 0; x < 5; x++)
			{
				double e = 1;
				if (x == 0)
					e = 4;
				if (x == 4)
					e = 5;
				if (y == 0)
					e = 6;
				if (y == 4)
					e = 7;
				if (z == 0)
					e = 8;
				if (z == 4)
					e = 9;
				EXPECT_FLOAT_EQ(dens(m, x, y, z, 0), e);
				EXPECT_FLOAT_EQ(dens(m, x, y, z, 1), 1);
			}
}

static void agents_simple(index_t dims, bool internalized, bool recompute, bool from_team)
{
	point_t<index_t, 3> maxs = { 60, dims >= 2 ? 60 : 20, dims >= 3 ? 60 : 20 };
	cartesian_mesh mesh(dims, { 0, 0, 0 }, maxs, { 20, 20, 20 });
	auto m = default_microenv(mesh);
	m.diffusion_time_step = 0.01f;
	m.compute_internalized_substrates = internalized;
	auto a1 = m.agents->create_agent();
	auto a2 = m.agents->create_agent();
	auto a3 = m.agents->create_agent();
	set_default_agent_values(a1, 0, 1000, { 10, 10, 10 }, dims);
	set_default_agent_values(a2, 400, 1000, { 30, 30, 30 }, dims);
	set_default_agent_values(a3, 800, 1000, { 50, 50, 50 }, dims);
	auto& s = the_solver();
	s.initialize(m);
	auto step = [&](bool rc) {
		if (from_team)
		{
			// host call sites run inside an OpenMP team (src/main.cpp:71); the facade enqueues once
			s.store_data_to_solver(m);
#pragma omp parallel
			s.cell.simulate_secretion_and_uptake(m, rc);
			s.load_data_from_solver(m);
		}
		else
			runit(s, m, [&] { s.cell.simulate_secretion_and_uptake(m, rc); });
	};
	step(true);
	const index_t d1 = dims >= 2 ? 1 : 0, d2 = dims >= 2 ? 2 : 0;
	auto at = [&](index_t i, index_t sb) {
		const index_t x = i, y = i == 0 ? 0 : (i == 1 ? d1 : d2), z = dims == 3 ? y : 0;
		return dens(m, x, y, z, sb);
	};
	if (internalized)
	{
		EXPECT_FLOAT_EQ(a1->internalized_substrates()[0], -216004.000000);
		EXPECT_FLOAT_EQ(a2->internalized_substrates()[0], -1469060.631579);
		EXPECT_FLOAT_EQ(a3->internalized_substrates()[0], -2927715.703704);
		EXPECT_FLOAT_EQ(a1->internalized_substrates()[1], 0);
	}
	EXPECT_FLOAT_EQ(at(0, 0), 28.000500);
	EXPECT_FLOAT_EQ(at(1, 0), 184.632579);
	EXPECT_FLOAT_EQ(at(2, 0), 366.964463);
	EXPECT_FLOAT_EQ(at(0, 1), 1);
	step(recompute);
	if (internalized)
	{
		EXPECT_FLOAT_EQ(a1->internalized_substrates()[0], -216004.000000 + -157093.818182);
		EXPECT_FLOAT_EQ(a2->internalized_substrates()[0], -1469060.631579 + -618551.844632);
		EXPECT_FLOAT_EQ(a3->internalized_substrates()[0], -2927715.703704 + -867471.319407);
	}
	EXPECT_FLOAT_EQ(at(0, 0), 47.637227);
	EXPECT_FLOAT_EQ(at(1, 0), 261.951560);
	EXPECT_FLOAT_EQ(at(2, 0), 475.398378);
}

// agents sharing voxels (test/test.cpp:1129-1214): brute-force per-voxel expectations
static void agents_conflict_big(bool internalized, bool recompute)
{
	cartesian_mesh mesh(1, { 0, 0, 0 }, { 2000, 20, 20 }, { 20, 20, 20 });
	auto m = default_microenv(mesh);
	m.diffusion_time_step = 0.01f;
	m.compute_internalized_substrates = internalized;
	std::vector<agent*> agents;
	for (index_t i = 0; i < mesh.grid_shape[0]; i++)
		for (int j = 0; j < 50; j++)
		{
			agents.push_back(m.agents->create_agent());
			set_default_agent_values(agents.back(), 0, 500, mesh.voxel_center({ i, 0, 0 }), 1);
		}
	auto& s = the_solver();
	s.initialize(m);
	const double vv = 8000, dt = 0.01;
	std::vector<double> exp_i(agents.size(), 0.0);
	for (bool rc : { true, recompute })
	{
		std::vector<double> exp_d(100);
		for (index_t x = 0; x < 100; x++)
		{
			double num = 0, den = 0, fac = 0;
			for (int j = 0; j < 50; j++)
			{
				num += 100. * 300 * dt * 500 / vv;
				den += (100. + 200) * dt * 500 / vv;
				fac += 400. * dt / vv;
			}
			const double D = dens(m, x, 0, 0, 0);
			exp_d[x] = (D + num) / (1 + den) + fac;
			for (int j = 0; j < 50; j++)
			{
				const double n1 = 100. * 300 * dt * 500 / vv, d1 = (100. + 200) * dt * 500 / vv,
							 f1 = 400. * dt / vv;
				exp_i[x * 50 + j] -= vv * ((-d1 * D + n1) / (1 + d1) + f1);
			}
		}
		runit(s, m, [&] { s.cell.simulate_secretion_and_uptake(m, rc); });
		for (index_t x = 0; x < 100; x++)
		{
			EXPECT_FLOAT_EQ(dens(m, x, 0, 0, 0), exp_d[x]);
			EXPECT_FLOAT_EQ(dens(m, x, 0, 0, 1), 1);
		}
		if (internalized)
			for (std::size_t i = 0; i < agents.size(); i++)
				EXPECT_FLOAT_EQ(agents[i]->internalized_substrates()[0], exp_i[i]);
	}
}

static void builder_and_timestep()
{
	microenvironment_builder b;
	b.set_name("builder");
	b.set_time_step(5);
	b.resize(3, { 0, 0, 0 }, { 200, 200, 200 }, { 20, 20, 20 });
	b.add_density("a", "u", 4, 5, 1);
	b.add_density("b", "u", 2, 3, 1);
	bool threw = false;
	try
	{
		b.add_dirichlet_node({ 1, 1, 1 }, { 1 });
	}
	catch (const std::runtime_error&)
	{
		threw = true;
	}
	EXPECT_FLOAT_EQ(threw ? 1 : 0, 1);
	auto m = b.build();
	EXPECT_FLOAT_EQ(m.find_substrate_index("b"), 1);
	auto& s = the_solver();
	s.initialize(m);
	s.store_data_to_solver(m);
	s.timestep(m, true, true);
	s.load_data_from_solver(m);
	// a timestep without agents is one diffusion solve: the uniform closed form
	for (std::size_t v = 0; v < m.mesh.voxel_count(); v++)
	{
		EXPECT_FLOAT_EQ(m.substrate_densities[v * 2 + 0], 0.0012299563);
		EXPECT_FLOAT_EQ(m.substrate_densities[v * 2 + 1], 0.0046296306);
	}
	// uniform field: zero gradient
	for (std::size_t i = 0; i < m.mesh.voxel_count() * 2 * 3; i++)
		EXPECT_FLOAT_EQ(m.gradients[i] + 1.0f, 1.0f);
}

// bulk_solver through the facade, the callbacks sampled at initialize: per voxel, with S = x+1, U = y, T = z+2
// per substrate index s (scaled by s+1), n solves from D0 = 1 against the closed form of the recurrence
// D' = (D + dt*S*T) / (1 + dt*(U+S)) evaluated in double precision.
static void bulk_D3(bool from_team)
{
	auto m = default_microenv(cartesian_mesh(3, { 0, 0, 0 }, { 100, 80, 60 }, { 20, 20, 20 }));
	m.supply_rate_func = [](microenvironment& mm, point_t<index_t, 3> v, real_t* out) {
		for (index_t s = 0; s < mm.substrates_count; s++)
			out[s] = (real_t)(v[0] + 1) * (s + 1);
	};
	m.uptake_rate_func = [](microenvironment& mm, point_t<index_t, 3> v, real_t* out) {
		for (index_t s = 0; s < mm.substrates_count; s++)
			out[s] = (real_t)v[1] * (s + 1);
	};
	m.supply_target_densities_func = [](microenvironment& mm, point_t<index_t, 3> v, real_t* out) {
		for (index_t s = 0; s < mm.substrates_count; s++)
			out[s] = (real_t)(v[2] + 2);
	};
	auto& s = the_solver();
	s.initialize(m);
	const int n = 3;
	auto run = [&] {
		s.store_data_to_solver(m);
		for (int i = 0; i < n; i++)
			s.bulk.solve(m);
		s.load_data_from_solver(m);
	};
	if (from_team)
	{
#pragma omp parallel num_threads(4)
		run();
	}
	else
		run();
	const double dt = m.diffusion_time_step;
	for (index_t z = 0; z < m.mesh.grid_shape[2]; z++)
		for (index_t y = 0; y < m.mesh.grid_shape[1]; y++)
			for (index_t x = 0; x < m.mesh.grid_shape[0]; x++)
				for (index_t sb = 0; sb < 2; sb++)
				{
					const double S = (x + 1.0) * (sb + 1), U = (double)y * (sb + 1), T = z + 2.0;
					double d = 1;
					for (int i = 0; i < n; i++)
						d = (d + dt * S * T) / (1 + dt * (U + S));
					EXPECT_FLOAT_EQ(dens(m, x, y, z, sb), d);
				}
}

// Agent lifecycle across the boundary (include/BioFVM/agent_container_base.h:61-92, src/agent_data.cpp:27-53,
// CUDA/cell_solver.cpp:180-211): agents one per voxel, a secretion step, a swap-remove from the middle, growth
// past the device allocation, an agent moved to another voxel, then steps with and without recompute.  Every
// voxel is checked against the single-agent closed form (host/cell_solver.h:383-401) in double precision:
// D' = (D + num) / (1 + den) + fac and I' = I - vv ((num - den D) / (1 + den) + fac), with num = S T dt vol / vv,
// den = (U + S) dt vol / vv, fac = N dt / vv; voxels without an agent keep their density.
static void agents_lifecycle()
{
	cartesian_mesh mesh(3, { 0, 0, 0 }, { 120, 100, 80 }, { 20, 20, 20 }); // 6 x 5 x 4 voxels
	auto m = default_microenv(mesh);
	m.diffusion_time_step = 0.01f;
	m.compute_internalized_substrates = true;
	const index_t S = m.substrates_count;
	const double dt = m.diffusion_time_step, vv = 8000;
	for (std::size_t i = 0; i < m.mesh.voxel_count() * S; i++)
		m.substrate_densities[i] = 0.5f + 0.001f * (real_t)(i % 97);

	// agent k sits in voxel `cell` (x, y, z); its rates follow from k
	auto place = [&](agent* a, index_t k, index_t cell) {
		const index_t x = cell % 6, y = (cell / 6) % 5, z = cell / 30;
		for (index_t sb = 0; sb < S; sb++)
		{
			a->secretion_rates()[sb] = 1.0f + 0.25f * (real_t)k + (real_t)sb;
			a->uptake_rates()[sb] = 0.5f + 0.125f * (real_t)k;
			a->saturation_densities()[sb] = 2.0f + 0.5f * (real_t)sb;
			a->net_export_rates()[sb] = 10.0f * (real_t)(k % 3);
			a->fraction_released_at_death()[sb] = 0.5f;
		}
		a->volume() = 1000 + 10 * (real_t)k;
		a->position()[0] = 20 * (real_t)x + 7;
		a->position()[1] = 20 * (real_t)y + 11;
		a->position()[2] = 20 * (real_t)z + 13;
	};
	auto voxel_of = [&](agent* a) {
		const auto v = a->voxel_index();
		return ((std::size_t)v[2] * 5 + v[1]) * 6 + v[0];
	};
	std::vector<agent*> live;
	index_t next_k = 0, next_cell = 0;
	auto make = [&](int count) {
		for (int i = 0; i < count; i++)
		{
			agent* a = m.agents->create_agent();
			place(a, next_k++, next_cell);
			next_cell += 7; // distinct voxels: 7 is coprime with the 120 voxels
			next_cell %= 120;
			live.push_back(a);
		}
	};
	auto& s = the_solver();
	// per agent and substrate: magnitudes of the terms its internalized amount accumulated
	std::map<const agent*, std::vector<double>> term_scale;
	// one step and its closed-form check over every voxel
	auto step_and_check = [&](bool recompute) {
		std::vector<double> before(m.substrate_densities.get(), m.substrate_densities.get() + m.mesh.voxel_count() * S);
		std::vector<double> before_I;
		for (agent* a : live)
			for (index_t sb = 0; sb < S; sb++)
				before_I.push_back(a->internalized_substrates()[sb]);

		s.store_data_to_solver(m);
		s.cell.simulate_secretion_and_uptake(m, recompute);
		s.load_data_from_solver(m);
		std::vector<double> expect = before;
		for (std::size_t i = 0; i < live.size(); i++)
		{
			agent* a = live[i];
			const std::size_t v = voxel_of(a);
			for (index_t sb = 0; sb < S; sb++)
			{
				const double Sr = a->secretion_rates()[sb], U = a->uptake_rates()[sb], T = a->saturation_densities()[sb];
				const double N = a->net_export_rates()[sb], vol = a->volume();
				const double num = Sr * T * dt * vol / vv, den = (U + Sr) * dt * vol / vv, fac = N * dt / vv;
				const double D = before[v * S + sb];
				expect[v * S + sb] = (D + num) / (1 + den) + fac;
				const double I = before_I[i * S + sb] - vv * ((num - den * D) / (1 + den) + fac);
				auto& scales = term_scale[a];
				scales.resize(S, 0.0);
				double& scale = scales[sb];
				scale += vv * (std::fabs(num) + std::fabs(den * D) + std::fabs(fac)) + std::fabs(before_I[i * S + sb]);
				expect_near_scaled(a->internalized_substrates()[sb], I, scale, __FILE__, __LINE__);
			}
		}
		for (std::size_t i = 0; i < expect.size(); i++)
			EXPECT_FLOAT_EQ(m.substrate_densities[i], expect[i]);
	};

	make(6);
	s.initialize(m);
	step_and_check(true);

	// swap-remove from the middle: the last agent takes slot 2, its handle and data follow
	agent* last = live.back();
	const real_t last_secretion = last->secretion_rates()[1];
	m.agents->remove_at(2);
	live.erase(live.begin() + 2);
	live.pop_back();
	live.insert(live.begin() + 2, last);
	EXPECT_FLOAT_EQ(m.agents->agents_count(), 5);
	EXPECT_FLOAT_EQ(last->index(), 2);
	EXPECT_FLOAT_EQ(m.agents->get_agent_at(2)->secretion_rates()[1], last_secretion);
	EXPECT_FLOAT_EQ(m.agents->get_agent_at(2) == last ? 1 : 0, 1);
	step_and_check(true);

	// growth past the device allocation (6 agents) and an agent moved to a voxel no other agent uses
	make(14);
	live[0]->position()[0] = 20 * 1 + 3;
	live[0]->position()[1] = 3;
	live[0]->position()[2] = 3; // voxel (1, 0, 0): 7k mod 120 for k < 20 never gives 1
	bool clash = false;
	for (std::size_t i = 1; i < live.size(); i++)
		clash |= voxel_of(live[i]) == voxel_of(live[0]);
	EXPECT_FLOAT_EQ(clash ? 1 : 0, 0);
	step_and_check(true);
	// the reduced factors are reused without recompute
	step_and_check(false);

	// an agent removed by handle (the last slot: no move), then a step
	agent* gone = live.back();
	m.agents->remove_agent(gone);
	live.pop_back();
	EXPECT_FLOAT_EQ(m.agents->agents_count(), (double)live.size());
	step_and_check(true);

	// out-of-range slots are rejected
	bool threw = false;
	try
	{
		m.agents->get_agent_at((index_t)live.size());
	}
	catch (const std::out_of_range&)
	{
		threw = true;
	}
	EXPECT_FLOAT_EQ(threw ? 1 : 0, 1);
}

// The builder's bookkeeping (host only, no device work): densities by name, faces created only where some
// substrate has a condition, interior nodes flattened in insertion order with default conditions.
static void builder_layout()
{
	microenvironment_builder b;
	b.resize(2, { 0, 0, 0 }, { 100, 60, 0 }, { 20, 20, 20 });
	b.add_density("o2", "mmHg", 100, 0.1f, 38);
	b.add_density("glc", "mM", 50, 0.2f, 5);
	b.add_density("lac", "mM", 25, 0.3f, 0);
	EXPECT_FLOAT_EQ(b.get_density_index("glc"), 1);
	bool threw = false;
	try
	{
		b.get_density_index("none");
	}
	catch (const std::runtime_error&)
	{
		threw = true;
	}
	EXPECT_FLOAT_EQ(threw ? 1 : 0, 1);
	b.add_boundary_dirichlet_conditions(1, { 1, 2, 3 }, { 4, 5, 6 }, { true, false, false }, { false, false, false });
	b.add_dirichlet_node({ 2, 1, 0 }, { 7, 8, 9 });
	b.add_dirichlet_node({ 4, 2, 0 }, { 10, 11, 12 }, { false, true, false });
	threw = false;
	try
	{
		b.add_boundary_dirichlet_conditions(3, { 0, 0, 0 }, { 0, 0, 0 });
	}
	catch (const std::runtime_error&)
	{
		threw = true;
	}
	EXPECT_FLOAT_EQ(threw ? 1 : 0, 1);
	b.do_compute_internalized_substrates();
	auto m = b.build();
	EXPECT_FLOAT_EQ(m.substrates_count, 3);
	EXPECT_FLOAT_EQ(m.mesh.grid_shape[0], 5);
	EXPECT_FLOAT_EQ(m.mesh.grid_shape[1], 3);
	EXPECT_FLOAT_EQ(m.diffusion_coefficients[2], 25);
	EXPECT_FLOAT_EQ(m.decay_rates[1], 0.2f);
	EXPECT_FLOAT_EQ(m.substrate_densities[0], 38);
	EXPECT_FLOAT_EQ(m.substrate_densities[1], 5);
	EXPECT_FLOAT_EQ(m.compute_internalized_substrates ? 1 : 0, 1);
	// only the x- face has a condition (substrate 1)
	EXPECT_FLOAT_EQ(m.dirichlet_min_boundary_values[0] ? 1 : 0, 1);
	EXPECT_FLOAT_EQ(m.dirichlet_min_boundary_values[1] ? 1 : 0, 0);
	EXPECT_FLOAT_EQ(m.dirichlet_max_boundary_values[0] ? 1 : 0, 0);
	EXPECT_FLOAT_EQ(m.dirichlet_min_boundary_values[0][1], 1);
	EXPECT_FLOAT_EQ(m.dirichlet_min_boundary_conditions[0][1] ? 1 : 0, 1);
	EXPECT_FLOAT_EQ(m.dirichlet_min_boundary_conditions[0][0] ? 1 : 0, 0);
	EXPECT_FLOAT_EQ(m.dirichlet_interior_voxels_count, 2);
	const index_t want_vox[4] = { 2, 1, 4, 2 };
	for (int i = 0; i < 4; i++)
		EXPECT_FLOAT_EQ(m.dirichlet_interior_voxels[i], want_vox[i]);
	for (int i = 0; i < 6; i++)
		EXPECT_FLOAT_EQ(m.dirichlet_interior_values[i], 7 + i);
	const bool want_cond[6] = { true, true, true, false, true, false };
	for (int i = 0; i < 6; i++)
		EXPECT_FLOAT_EQ(m.dirichlet_interior_conditions[i] ? 1 : 0, want_cond[i] ? 1 : 0);
}

// The container's slot bookkeeping (host only): swap-remove keeps handles and SoA slots in step, a new agent
// after a removal starts zeroed.
static void agent_container_slots()
{
	auto m = default_microenv(cartesian_mesh(2, { 0, 0, 0 }, { 100, 100, 0 }, { 20, 20, 20 }));
	std::vector<agent*> made;
	for (int i = 0; i < 5; i++)
	{
		agent* a = m.agents->create_agent();
		a->secretion_rates()[0] = (real_t)i;
		a->internalized_substrates()[1] = 100 + (real_t)i;
		a->volume() = 10 + (real_t)i;
		a->position()[1] = 1000 + (real_t)i;
		made.push_back(a);
	}
	EXPECT_FLOAT_EQ(made[3]->index(), 3);
	m.agents->remove_agent(made[1]); // agent 4 moves into slot 1
	EXPECT_FLOAT_EQ(m.agents->agents_count(), 4);
	EXPECT_FLOAT_EQ(made[4]->index(), 1);
	EXPECT_FLOAT_EQ(made[4]->secretion_rates()[0], 4);
	EXPECT_FLOAT_EQ(made[4]->internalized_substrates()[1], 104);
	EXPECT_FLOAT_EQ(made[4]->volume(), 14);
	EXPECT_FLOAT_EQ(made[4]->position()[1], 1004);
	EXPECT_FLOAT_EQ(m.agents->get_agent_at(1) == made[4] ? 1 : 0, 1);
	m.agents->remove_at(3); // the last slot: nothing moves
	EXPECT_FLOAT_EQ(m.agents->agents_count(), 3);
	EXPECT_FLOAT_EQ(made[2]->index(), 2);
	agent* fresh = m.agents->create_agent();
	EXPECT_FLOAT_EQ(fresh->index(), 3);
	EXPECT_FLOAT_EQ(fresh->secretion_rates()[0], 0);
	EXPECT_FLOAT_EQ(fresh->internalized_substrates()[1], 0);
	EXPECT_FLOAT_EQ(fresh->volume(), 0);
	EXPECT_FLOAT_EQ(fresh->position()[1], 0);
	EXPECT_FLOAT_EQ(m.agents->agents_count(), 4);
}

struct test_case
{
	const char* name;
	std::function<void()> run;
};

int main(int argc, char** argv)
{
	const double D3_random[] = {
		0.6333066643,  1.6268066007,  2.5825920996,	 3.5703051208,	4.5318775349,  5.5138036408,  6.4811629703,
		7.4573021609,  8.4304484056,  9.4008006809,	 10.3797338410, 11.3442992010, 12.3290192763, 13.2877977210,
		14.2783047117, 15.2312962410, 16.2275901470, 17.1747947611, 18.1768755823, 19.1182932811, 20.1261610177,
		21.0617918012, 22.0754464530, 23.0052903212, 24.0247318884, 24.9487888412, 25.9740173237, 26.8922873613,
		27.9233027591, 28.8357858813, 29.8725881944, 30.7792844014, 31.8218736297, 32.7227829214, 33.7711590651,
		34.6662814414, 35.7204445004, 36.6097799615, 37.6697299358, 38.5532784815, 39.6190153711, 40.4967770016,
		41.5683008064, 42.4402755216, 43.5175862418, 44.3837740416, 45.4668716771, 46.3272725617, 47.4161571125,
		48.2707710817, 49.3654425478, 50.2142696018, 51.3147279832, 52.1577681218
	};
	const double D2_random[] = { 0.1948319355,	1.1899772978,  2.1441254507,  3.1335099015,	 4.0934189658,
								 5.0770425053,	6.0427124809,  7.0205751090,  7.9920082,	 8.9641077127,
								 9.9412995111,	10.9076403164, 11.8905930262, 12.8511729202, 13.8398865413,
								 14.7947055239, 15.7891800565, 16.7382381276 };
	std::vector<test_case> cases = {
		{ "diffusion_D2_uniform_200", [] { diffusion_uniform(2, { 200, 200, 0 }, { 20, 20, 0 }, 0.0054869675, 0.013840831); } },
		{ "diffusion_D2_uniform_20000", [] { diffusion_uniform(2, { 20000, 20000, 0 }, { 20, 20, 0 }, 0.0054869675, 0.013840831); } },
		{ "diffusion_D3_uniform", [] { diffusion_uniform(3, { 200, 200, 200 }, { 20, 20, 20 }, 0.0012299563, 0.0046296306); } },
		{ "diffusion_D2_random", [&] { diffusion_random(2, { 60, 60, 0 }, std::vector<double>(D2_random, D2_random + 18)); } },
		{ "diffusion_D3_random", [&] { diffusion_random(3, { 60, 60, 60 }, std::vector<double>(D3_random, D3_random + 54)); } },
		{ "dirichlet_one_cond_D3", dirichlet_one_cond_D3 },
		{ "dirichlet_multiple_cond_D2", dirichlet_multiple_cond_D2 },
		{ "gradient_D3", gradient_D3 },
		{ "dirichlet_boundaries_D3", boundaries_D3 },
		{ "builder_and_timestep", builder_and_timestep },
		{ "bulk_D3", [] { bulk_D3(false); } },
		{ "bulk_D3_from_omp_team", [] { bulk_D3(true); } },
		{ "agents_lifecycle", agents_lifecycle },
		{ "host_builder_layout", builder_layout },
		{ "host_agent_container_slots", agent_container_slots },
	};
	for (index_t dims : { 1, 2, 3 })
		for (bool ci : { true, false })
			for (bool rc : { true, false })
				cases.push_back({ "agents_simple", [=] { agents_simple(dims, ci, rc, false); } });
	cases.push_back({ "agents_simple_from_omp_team", [] { agents_simple(3, true, false, true); } });
	for (bool ci : { true, false })
		for (bool rc : { true, false })
			cases.push_back({ "agents_conflict_big", [=] { agents_conflict_big(ci, rc); } });

	const char* only = argc > 1 ? argv[1] : nullptr;
	int failed_cases = 0, run = 0;
	for (auto& c : cases)
	{
		if (only && std::strstr(c.name, only) == nullptr)
			continue;
		g_test = c.name;
		const int before = g_failures;
		std::printf("[ RUN  ] %s\n", c.name);
		std::fflush(stdout);
		try
		{
			c.run();
		}
		catch (const std::exception& e)
		{
			std::printf("  EXCEPTION %s: %s\n", c.name, e.what());
			g_failures++;
		}
		const bool ok = g_failures == before;
		failed_cases += ok ? 0 : 1;
		run++;
		std::printf("[ %s ] %s\n", ok ? " OK " : "FAIL", c.name);
		std::fflush(stdout);
	}
	std::printf("%d cases, %d checks, %d failed checks, %d failed cases\n", run, g_checks, g_failures, failed_cases);
	return failed_cases == 0 ? 0 : 1;
}
