#include "agent_data.h"

#include "data_utils.h"
#include "microenvironment.h"

using namespace biofvm;

agent_data::agent_data(microenvironment& m) : agents_count(0), m(m) {}

// src/agent_data.cpp:10-25 of the reference: grow every array by one agent, zero-filled
void agent_data::add()
{
	agents_count++;
	const std::size_t per_s = (std::size_t)agents_count * m.substrates_count;
	for (auto* v : { &secretion_rates, &saturation_densities, &uptake_rates, &net_export_rates,
					 &internalized_substrates, &fraction_released_at_death, &fraction_transferred_when_ingested })
		v->resize(per_s, 0);
	volumes.resize(agents_count, 0);
	positions.resize((std::size_t)agents_count * m.mesh.dims, 0);
}

// src/agent_data.cpp:27-53 of the reference: the last agent moves into the removed slot
void agent_data::remove(index_t index)
{
	agents_count--;
	if (index == agents_count)
		return;
	const index_t S = m.substrates_count;
	for (auto* v : { &secretion_rates, &saturation_densities, &uptake_rates, &net_export_rates,
					 &internalized_substrates, &fraction_released_at_death, &fraction_transferred_when_ingested })
		move_vector(v->data() + (std::size_t)index * S, v->data() + (std::size_t)agents_count * S, S);
	volumes[index] = volumes[agents_count];
	move_vector(positions.data() + (std::size_t)index * m.mesh.dims,
				positions.data() + (std::size_t)agents_count * m.mesh.dims, m.mesh.dims);
}
