// Wraps any frontier-like type (clear/push/pop/size/empty) as a SEPFrontier
// (interface of the original include/sep/sep_frontier_adapter.hpp:11-104).
#pragma once

#include "sep_frontier.hpp"

#include <cstddef>

namespace hytgraph::sep
{
    template <typename Frontier>
    class SEPFrontierAdapter final : public SEPFrontier
    {
    public:
        using node_id_type = SEPFrontier::node_id_type;
        using frontier_type = Frontier;

        explicit SEPFrontierAdapter(Frontier &frontier) noexcept : frontier_(&frontier) {}

        void clear() noexcept override { frontier_->clear(); }
        void push(node_id_type node) override { frontier_->push(node); }
        bool pop(node_id_type &node) override { return frontier_->pop(node); }
        std::size_t size() const noexcept override { return frontier_->size(); }
        bool empty() const noexcept override { return frontier_->empty(); }
        bool is_deterministic() const noexcept override { return false; }

        Frontier &underlying() noexcept { return *frontier_; }
        const Frontier &underlying() const noexcept { return *frontier_; }

    private:
        Frontier *frontier_;
    };

    template <typename Frontier>
    SEPFrontierAdapter<Frontier> make_sep_frontier_adapter(Frontier &frontier) noexcept
    {
        return SEPFrontierAdapter<Frontier>(frontier);
    }
} // namespace hytgraph::sep
