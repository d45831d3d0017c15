// Agent containers with the reference's public interface (include/BioFVM/agent_container_base.h:9-107):
// agents live in a structure-of-arrays (agent_data) and are addressed through stable `agent` handles.  A new
// agent takes the slot after the last one; removing an agent moves the last agent's data and handle into the
// freed slot, so the arrays stay dense and a removal costs O(S).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
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

	// handles[i] is the agent whose data sit in slot i of data_
	std::vector<std::unique_ptr<agent_t>> agents_;

	template <typename... args_t>
	agent_container_common(args_t&&... args) : data_(std::forward<args_t>(args)...)
	{}

	void check_slot(index_t index) const
	{
		if (index < 0 || index >= (index_t)agents_.size())
			throw std::out_of_range("agent index " + std::to_string(index) + " out of range [0, " +
									std::to_string(agents_.size()) + ")");
	}

	template <typename wanted_t>
	static wanted_t* as(agent_t* a)
	{
		if constexpr (std::is_same_v<wanted_t, agent_t>)
			return a;
		else
			return dynamic_cast<wanted_t*>(a);
	}

public:
	virtual index_t agents_count() const override { return (index_t)agents_.size(); }

	virtual agent* create_agent() override { return create(); }

	template <typename other_agent_t = agent_t>
	other_agent_t* create()
	{
		const index_t slot = (index_t)agents_.size();
		data_.add();
		agents_.push_back(std::make_unique<other_agent_t>(next_agent_id_++, data_, slot));
		return as<other_agent_t>(agents_.back().get());
	}

	virtual void remove_agent(agent* a) override { remove_at(get_agent_index(a)); }

	virtual void remove_at(index_t index) override
	{
		check_slot(index);
		const index_t last = (index_t)agents_.size() - 1;
		if (index != last)
		{
			// the removed handle is destroyed here; the last agent now answers for `index`
			agents_[index] = std::move(agents_[last]);
			get_agent_index(agents_[index].get()) = index;
		}
		agents_.pop_back();
		data_.remove(index);
	}

	virtual agent* get_agent_at(index_t index) override { return get_at(index); }

	template <typename other_agent_t = agent_t>
	other_agent_t* get_at(index_t index)
	{
		check_slot(index);
		return as<other_agent_t>(agents_[index].get());
	}

	const std::vector<std::unique_ptr<agent_t>>& agents() const { return agents_; }

	const agent_data_t& data() const { return data_; }
};

} // namespace biofvm
