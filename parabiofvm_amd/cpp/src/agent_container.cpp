#include "agent_container.h"

using namespace biofvm;

agent_container_base::agent_container_base() : next_agent_id_(0) {}

index_t& agent_container_base::get_agent_index(agent* a) { return a->index_; }

agent_data& agent_container::get_agent_data() { return data_; }

agent_container::agent_container(microenvironment& m) : agent_container_common<agent, agent_data>(m) {}
