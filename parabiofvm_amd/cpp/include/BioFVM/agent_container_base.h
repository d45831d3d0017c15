// Agent containers (include/BioFVM/agent_container_base.h:9-107 of the reference): agents are created at the
// end of the SoA and removed by moving the last agent into the freed slot.
#pragma once

#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

#include "agent.h"
#include "agent_data.h"

namespace biofvm {

namespace solvers {
class common_solver;
}

using agent_ptr = std::unique_ptr<agent>;

class agent_container_base;

using agent_container_ptr = std::unique_ptr<agent_container_base>;

class agent_container_base
{
	friend solvers::common_solver;

protected:
	virtual agent_data& get_agent_data() = 0;

	agent_id_t next_agent_id_;

	agent_container_base();

	index_t& get_agent_index(agent* agent);

public:
	virtual index_t agents_count() const = 0;

	virtual agent* create_agent() = 0;

	virtual void remove_agent(agent* agent) = 0;
	virtual void remove_at(index_t index) = 0;

	virtual agent* get_agent_at(index_t index) = 0;

	virtual ~agent_container_base() = default;
};

template <typename agent_t, typename agent_data_t>
class agent_container_common : public agent_container_base
{
protected:
	agent_data_t data_;

	std::vector<std::unique_ptr<agent_t>> agents_;

	template <typename... args_t>
	agent_container_common(args_t&&... args) : data_(std::forward<args_t>(args)...)
	{}

public:
	virtual index_t agents_count() const override { return (index_t)agents_.size(); }

	virtual agent* create_agent() override { return create(); }

	template <typename other_agent_t = agent_t>
	other_agent_t* create()
	{
		data_.add();
		const index_t slot = data_.agents_count - 1;
		agents_.push_back(std::make_unique<other_agent_t>(next_agent_id_++, data_, slot));
		agent_t* made = agents_.back().get();
		if constexpr (std::is_same_v<other_agent_t, agent_t>)
			return made;
		else
			return dynamic_cast<other_agent_t*>(made);
	}

	virtual void remove_agent(agent* a) override { remove_at(get_agent_index(a)); }

	virtual void remove_at(index_t index) override
	{
		// the last agent takes the freed slot, in the handle list and in the SoA
		get_agent_index(agents_.back().get()) = index;
		std::swap(agents_[index], agents_.back());
		agents_.pop_back();
		data_.remove(index);
	}

	virtual agent* get_agent_at(index_t index) override { return get_at(index); }

	template <typename other_agent_t = agent_t>
	other_agent_t* get_at(index_t index)
	{
		if constexpr (std::is_same_v<other_agent_t, agent_t>)
			return agents_[index].get();
		else
			return dynamic_cast<other_agent_t*>(agents_[index].get());
	}

	const std::vector<std::unique_ptr<agent_t>>& agents() const { return agents_; }

	const agent_data_t& data() const { return data_; }
};

} // namespace biofvm
