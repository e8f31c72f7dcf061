// SEP applications and drivers bound to the MI355X kernels (no counterpart
// in the original, whose factory returns deferred drivers only; this is the
// concrete execution its SEP contract leaves open).
//
// SEPSSSPApplication / SEPPageRankApplication implement the SEP hooks with
// the min-plus (SSSP) and delta-push (PageRank) semantics, so host code can
// drive them through any SEP driver; SEPSSSPGpuDriver / SEPPageRankGpuDriver
// run the same computation on the GPU, one round per execute_step() (the
// hyt_stepper_* rounds of the C ABI): SUCCESS with the round's frontier rows
// (vertices_processed) and improvements (updates_generated) while work is
// left, CONVERGED on the round that finds none. The context's graph handle is
// a `const graph::CSRGraph*`; the device graph is uploaded on the first step
// and kept until the driver is finalised or sees another graph. values() is
// the state after the last step (distances, or normalised ranks).
// SEPExecutionFactory hands these drivers out for these applications.
//
// run_to_convergence() executes every remaining SSSP round at once; under
// the DATA_DRIVEN PUSH variants it is one persistent kernel launch.
//
// Variant mapping (results are the same for every variant; the variant picks
// how rounds run):
//   SYNC  / ASYNC -> synchronous rounds (the original's Jacobi semantics and
//                    iteration count) / rounds reading live values (SSSP;
//                    PageRank always pushes residuals as they arrive);
//   PUSH  / PULL  -> push along out-edges / pull along in-edges;
//   DATA_DRIVEN   -> frontier rounds;
//   TOPOLOGY_DRIVEN -> every round is a pull over all vertices.
#pragma once

#include "graph/csr_graph.hpp"
#include "sep_execution_driver_base.hpp"

#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <vector>

namespace hytgraph::sep
{
    class SEPSSSPApplication final : public SEPApplication<float, float, float>
    {
    public:
        explicit SEPSSSPApplication(node_id_type source) noexcept : source_(source) {}

        node_id_type source() const noexcept { return source_; }

        void set_execution_variant(const SEPExecutionVariant &v) override { variant_ = v; }
        const SEPExecutionVariant &execution_variant() const noexcept override { return variant_; }
        float init_value(node_id_type n) const override { return n == source_ ? 0.0F : kInf; }
        float init_buffer(node_id_type n) const override { return init_value(n); }
        float identity_element() const override { return kInf; }
        CombineResult combine_value_buffer(node_id_type, const float &value, const float &buffer) const override
        {
            return buffer < value ? CombineResult{buffer, true} : CombineResult{value, false};
        }
        AccumulateResult accumulate_buffer(node_id_type s, node_id_type d, float &dst, const float &msg) const override
        {
            return accumulate_buffer(s, d, 1.0F, dst, msg);
        }
        AccumulateResult accumulate_buffer(node_id_type, node_id_type, float w, float &dst,
                                           const float &msg) const override
        {
            const float cand = msg + w;
            if (!(cand < dst))
                return {};
            dst = cand;
            return {true, true};
        }
        bool is_active(node_id_type, const float &buffer) const override { return buffer < kInf; }
        bool is_high_priority(const float &current, const float &buffer) const override { return buffer < current; }
        bool requires_edge_weights() const noexcept override { return true; }

    private:
        static constexpr float kInf = std::numeric_limits<float>::infinity();
        node_id_type source_;
        SEPExecutionVariant variant_{ExecutionMode::SYNC, MessagePassing::PUSH, SchedulingMode::DATA_DRIVEN};
    };

    class SEPPageRankApplication final : public SEPApplication<float, float, double>
    {
    public:
        SEPPageRankApplication(float damping, float epsilon) noexcept : damping_(damping), epsilon_(epsilon) {}

        float damping() const noexcept { return damping_; }
        float epsilon() const noexcept { return epsilon_; }

        void set_execution_variant(const SEPExecutionVariant &v) override { variant_ = v; }
        const SEPExecutionVariant &execution_variant() const noexcept override { return variant_; }
        // value = accumulated rank, buffer = residual still to push
        float init_value(node_id_type) const override { return 0.0F; }
        float init_buffer(node_id_type) const override { return 1.0F - damping_; }
        float identity_element() const override { return 0.0F; }
        CombineResult combine_value_buffer(node_id_type, const float &, const float &buffer) const override
        {
            return {0.0F, buffer > epsilon_};
        }
        AccumulateResult accumulate_buffer(node_id_type, node_id_type, float &dst, const float &msg) const override
        {
            const bool was = dst > epsilon_;
            dst += msg;
            return {msg != 0.0F, !was && dst > epsilon_};
        }
        bool is_active(node_id_type, const float &buffer) const override { return buffer > epsilon_; }
        bool is_high_priority(const float &current, const float &buffer) const override { return buffer > current; }

    private:
        float damping_, epsilon_;
        SEPExecutionVariant variant_{ExecutionMode::SYNC, MessagePassing::PUSH, SchedulingMode::DATA_DRIVEN};
    };

    // A device graph and the round-by-round run on it (defined in the
    // library; opaque here).
    struct SEPGpuRun;

    // Shared part of the GPU drivers: variant checks, context access, the
    // run kept across steps, results.
    template <typename TWeight>
    class SEPGpuDriverBase : public SEPExecutionDriverBase<float, float, TWeight>
    {
    public:
        using base_type = SEPExecutionDriverBase<float, float, TWeight>;
        using step_result_type = typename base_type::step_result_type;

        void set_context(SEPExecutionContext &context) noexcept override { base_type::set_context(context); }
        void finalize() noexcept override
        {
            run_.reset();
            base_type::finalize();
        }

        // Per-vertex state after the last execute_step (distances or ranks).
        const std::vector<float> &values() const noexcept { return values_; }
        // Rounds run so far (SYNC SSSP: the original's iteration count).
        std::uint64_t rounds() const noexcept { return rounds_; }

    protected:
        const graph::CSRGraph *graph_or_null()
        {
            return this->context_is_valid() ? static_cast<const graph::CSRGraph *>(this->context().graph())
                                            : nullptr;
        }

        std::shared_ptr<SEPGpuRun> run_;
        std::vector<float> values_;
        std::uint64_t rounds_ = 0;
    };

    class SEPSSSPGpuDriver final : public SEPGpuDriverBase<float>
    {
    public:
        void initialize(application_type &application, const execution_config_type &config) override;
        step_result_type execute_step() override;
        // Every remaining round in one call (no counterpart in the original's
        // driver interface). A run not yet stepped under a DATA_DRIVEN PUSH
        // variant (SYNC_PUSH_DD, ASYNC_PUSH_DD) executes as one persistent
        // cooperative launch (hyt_run_options.persistent): the next queue is
        // built on the device and no round returns to the host; SYNC keeps
        // the original's iteration count (rounds()). Other variants, or a
        // run already stepped, take the remaining rounds through the stepper.
        // CONVERGED with the run's frontier rows and improvements.
        step_result_type run_to_convergence();
        // True when the last run_to_convergence() ran the persistent kernel.
        bool ran_persistent() const noexcept { return persistent_; }

    private:
        bool persistent_ = false;
    };

    class SEPPageRankGpuDriver final : public SEPGpuDriverBase<double>
    {
    public:
        void initialize(application_type &application, const execution_config_type &config) override;
        step_result_type execute_step() override;
    };
} // namespace hytgraph::sep
