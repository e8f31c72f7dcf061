// Abstract SEP frontier and a non-owning view of one (interface of the
// original include/sep/sep_frontier.hpp:11-135).
#pragma once

#include <cstddef>
#include <vector>

namespace hytgraph::sep
{
    class SEPFrontier
    {
    public:
        using node_id_type = std::size_t;

        virtual ~SEPFrontier() = default;
        virtual void clear() noexcept = 0;
        virtual void push(node_id_type node) = 0;
        virtual void push(const std::vector<node_id_type> &nodes)
        {
            for (node_id_type n : nodes)
                push(n);
        }
        virtual bool pop(node_id_type &node) = 0;
        virtual std::size_t size() const noexcept = 0;
        virtual bool empty() const noexcept = 0;
        virtual bool is_deterministic() const noexcept { return false; }
    };

    class SEPFrontierView
    {
    public:
        explicit SEPFrontierView(SEPFrontier *frontier) noexcept : frontier_(frontier) {}

        SEPFrontier *get() noexcept { return frontier_; }
        const SEPFrontier *get() const noexcept { return frontier_; }
        bool valid() const noexcept { return frontier_ != nullptr; }
        std::size_t size() const noexcept { return frontier_ ? frontier_->size() : 0; }
        bool empty() const noexcept { return frontier_ == nullptr || frontier_->empty(); }

    private:
        SEPFrontier *frontier_;
    };
} // namespace hytgraph::sep
