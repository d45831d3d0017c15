#include "agent_data.h"

#include "data_utils.h"
#include "microenvironment.h"

using namespace biofvm;

agent_data::agent_data(microenvironment& m) : agents_count(0), m(m) {}

namespace {

// every per-(agent, substrate) array of the SoA
template <typename F>
void for_each_substrate_array(agent_data& d, F&& f)
{
	for (std::vector<real_t>* v : { &d.secretion_rates, &d.saturation_densities, &d.uptake_rates, &d.net_export_rates,
									&d.internalized_substrates, &d.fraction_released_at_death,
									&d.fraction_transferred_when_ingested })
		f(*v);
}

} // namespace

// The arrays always hold exactly agents_count agents, so growing them by one appends a zero-filled agent
// (the reference's add, src/agent_data.cpp:10-25, grows them the same way).
void agent_data::add()
{
	agents_count++;
	const std::size_t S = (std::size_t)m.substrates_count;
	for_each_substrate_array(*this, [&](std::vector<real_t>& v) { v.resize((std::size_t)agents_count * S, 0); });
	volumes.resize(agents_count, 0);
	positions.resize((std::size_t)agents_count * m.mesh.dims, 0);
}

// Swap-remove (src/agent_data.cpp:27-53): the last agent's data fill slot `index`, then every array drops its
// last agent.
void agent_data::remove(index_t index)
{
	const index_t last = agents_count - 1;
	const std::size_t S = (std::size_t)m.substrates_count;
	const std::size_t dims = (std::size_t)m.mesh.dims;
	if (index != last)
	{
		for_each_substrate_array(*this, [&](std::vector<real_t>& v) {
			move_vector(v.data() + (std::size_t)index * S, v.data() + (std::size_t)last * S, (index_t)S);
		});
		move_scalar(volumes.data() + index, volumes.data() + last);
		move_vector(positions.data() + (std::size_t)index * dims, positions.data() + (std::size_t)last * dims,
					(index_t)dims);
	}
	agents_count = last;
	for_each_substrate_array(*this, [&](std::vector<real_t>& v) { v.resize((std::size_t)last * S); });
	volumes.resize(last);
	positions.resize((std::size_t)last * dims);
}
