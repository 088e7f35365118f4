// FlowAnalyzer<2> with the reference's API (llalbm/src/utils/analysis/
// FlowAnalyzer.hpp:128-585).  Inside perform_lbm the lift/drag of the
// considered obstacle points is reduced on the GPU (lbm_lift_drag, FP64);
// compute_flow_properties on a host tensor evaluates the same formula on
// the CPU for post-hoc analysis.  CSV output matches the reference:
// ./flow_analysis/contribution_<k>.csv per analysis step and
// total_lift_drag.csv at the end.
#ifndef LLALBM_HIP_FLOW_ANALYZER_HPP
#define LLALBM_HIP_FLOW_ANALYZER_HPP

#include <cmath>
#include <filesystem>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/lbm_amd.h"
#include "aliases.hpp"

namespace llalbm::analysis
{
using util::HostTensor;
using util::Index;
using util::ObstaclePoint;

template <std::size_t dim>
class FlowAnalyzer;

template <>
class FlowAnalyzer<2> : public std::enable_shared_from_this<FlowAnalyzer<2>>
{
public:
    FlowAnalyzer(const std::size_t &iterations_between_save_ = 10) : iterations_between_save(iterations_between_save_)
    {
        initialize();
    }
    FlowAnalyzer(std::vector<ObstaclePoint<2>> &points, const std::size_t &iterations_between_save_ = 10,
                 const bool &should_save_ = true)
        : considered_points(points), iterations_between_save(iterations_between_save_), should_save(should_save_)
    {
        initialize();
    }
    static std::shared_ptr<FlowAnalyzer<2>> create(std::vector<ObstaclePoint<2>> &points,
                                                   const std::size_t &iterations_between_save_, const bool &should_save_)
    {
        return std::make_shared<FlowAnalyzer<2>>(points, iterations_between_save_, should_save_);
    }

    /// FlowAnalyzer.hpp:226-252: output directory, cleared of CSV files.
    void initialize()
    {
        save_counter = 0;
        if (!std::filesystem::create_directory("./flow_analysis"))
            for (const auto &entry : std::filesystem::directory_iterator("./flow_analysis"))
                if (entry.path().extension() == ".csv") std::filesystem::remove(entry.path());
    }

    /// FlowAnalyzer.hpp:259-268.
    void save_global_results() const
    {
        std::ofstream o("./flow_analysis/total_lift_drag.csv");
        o << "iteration,total_lift,total_drag" << std::endl;
        for (std::size_t i = 0; i < lift_drag_per_iteration.size(); ++i)
            o << i << "," << lift_drag_per_iteration[i].first << "," << lift_drag_per_iteration[i].second << std::endl;
    }

    void add_point(const ObstaclePoint<2> &p) { considered_points.push_back(p); }
    void add_point_vec(std::vector<ObstaclePoint<2>> &v) { considered_points = v; }
    void clear_points() { considered_points.clear(); }
    void set_iterations_between_save(const std::size_t &n) { iterations_between_save = n; }
    void set_should_save(const bool &s) { should_save = s; }
    std::size_t get_iterations_between_save() const { return iterations_between_save; }
    const std::vector<std::pair<double, double>> &get_lift_drag_history() const { return lift_drag_per_iteration; }

    /// FlowAnalyzer.hpp:348-451 on a host tensor of post-collision populations.
    std::pair<double, double> compute_flow_properties(const HostTensor<double, 3> &f)
    {
        std::vector<double> per(2 * considered_points.size());
        double lift = 0., drag = 0.;
        const double h = std::sqrt(2);
        for (std::size_t q = 0; q < considered_points.size(); ++q) {
            const auto &node = considered_points[q];
            const Index i = node.coords[0], j = node.coords[1];
            double ll = 0., ld = 0.;
            if (node.directions.test(1)) ld -= f(i, j + 1, 3);
            if (node.directions.test(2)) ll -= f(i - 1, j, 4);
            if (node.directions.test(3)) ld += f(i, j - 1, 1);
            if (node.directions.test(4)) ll += f(i + 1, j, 2);
            if (node.directions.test(5)) { ld -= (f(i - 1, j + 1, 7)) * h / 2.0; ll -= (f(i - 1, j + 1, 7)) * h / 2.0; }
            if (node.directions.test(6)) { ld += (f(i - 1, j - 1, 8)) * h / 2.0; ll -= (f(i - 1, j - 1, 8)) * h / 2.0; }
            if (node.directions.test(7)) { ld += (f(i + 1, j - 1, 5)) * h / 2.0; ll += (f(i + 1, j - 1, 5)) * h / 2.0; }
            if (node.directions.test(8)) { ld -= (f(i + 1, j + 1, 6)) * h / 2.0; ll += (f(i + 1, j + 1, 6)) * h / 2.0; }
            per[2 * q] = ll;
            per[2 * q + 1] = ld;
            drag += ld;
            lift += ll;
        }
        return record(lift, drag, per);
    }

    /// The same quantity reduced on the GPU from the post-collision
    /// populations of the context's last step (lbm_lift_drag).
    std::pair<double, double> compute_on_device(lbm_ctx *ctx)
    {
        collect();
        stage_points();
        double ld[2] = {0.0, 0.0};
        std::vector<double> per(2 * considered_points.size());
        if (!considered_points.empty()) {
            int r = lbm_lift_drag(ctx, ij.data(), masks.data(), (int64_t)considered_points.size(), ld,
                                  should_save ? per.data() : nullptr);
            if (r != LBM_OK) throw std::runtime_error(std::string("lbm_lift_drag: ") + lbm_last_error());
        }
        return record(ld[0], ld[1], per);
    }

    /// compute_on_device without waiting (lbm_lift_drag_async): the
    /// reduction queues behind the step on the GPU and its result lands in
    /// page-locked memory.  collect() records the results in call order;
    /// at most two are outstanding (a third call records the first).
    void compute_on_device_async(lbm_ctx *ctx)
    {
        Pending &p = pending[next];
        if (p.active) finish(p);
        p.seq = seq++;
        p.active = true;
        p.ctx = ctx;
        p.empty = considered_points.empty();
        p.n = considered_points.size();
        next ^= 1;
        if (p.empty) return;
        stage_points();
        const std::size_t n = considered_points.size();
        p.buf.reserve(2 + 2 * n);
        int r = lbm_lift_drag_async(ctx, ij.data(), masks.data(), (int64_t)n, p.buf.data(),
                                    should_save ? p.buf.data() + 2 : nullptr, &p.ticket);
        if (r != LBM_OK) {
            p.active = false;
            throw std::runtime_error(std::string("lbm_lift_drag_async: ") + lbm_last_error());
        }
    }

    /// Records every outstanding asynchronous result, oldest first.
    void collect()
    {
        Pending &a = pending[0], &b = pending[1];
        Pending *first = &a, *second = &b;
        if (b.active && (!a.active || b.seq < a.seq)) std::swap(first, second);
        if (first->active) finish(*first);
        if (second->active) finish(*second);
    }

private:
    /// Page-locked doubles for asynchronous results (lbm_host_alloc).
    struct PinnedDoubles {
        double *ptr = nullptr;
        std::size_t cap = 0;
        PinnedDoubles() = default;
        PinnedDoubles(const PinnedDoubles &) = delete;
        PinnedDoubles &operator=(const PinnedDoubles &) = delete;
        ~PinnedDoubles() { lbm_host_free(ptr); }
        void reserve(std::size_t n)
        {
            if (n <= cap) return;
            lbm_host_free(ptr);
            ptr = nullptr;
            cap = 0;
            void *p = nullptr;
            if (lbm_host_alloc((int64_t)(n * sizeof(double)), &p) != LBM_OK)
                throw std::runtime_error(std::string("lbm_host_alloc: ") + lbm_last_error());
            ptr = (double *)p;
            cap = n;
        }
        double *data() { return ptr; }
    };
    struct Pending {
        bool active = false, empty = false;
        int64_t ticket = 0, seq = 0;
        std::size_t n = 0;
        lbm_ctx *ctx = nullptr;
        PinnedDoubles buf;
    };

    void stage_points()
    {
        ij.resize(2 * considered_points.size());
        masks.resize(considered_points.size());
        for (std::size_t q = 0; q < considered_points.size(); ++q) {
            ij[2 * q] = considered_points[q].coords[0];
            ij[2 * q + 1] = considered_points[q].coords[1];
            masks[q] = (int32_t)considered_points[q].directions.to_ulong();
        }
    }

    void finish(Pending &p)
    {
        p.active = false;
        if (p.empty) {
            record(0.0, 0.0, {});
            return;
        }
        if (lbm_wait(p.ctx, p.ticket) != LBM_OK) throw std::runtime_error(std::string("lbm_wait: ") + lbm_last_error());
        const double *b = p.buf.data();
        std::vector<double> per;
        if (should_save) per.assign(b + 2, b + 2 + 2 * p.n);
        record(b[0], b[1], per);
    }

    std::pair<double, double> record(double lift, double drag, const std::vector<double> &per)
    {
        if (should_save) {
            std::ofstream o("./flow_analysis/contribution_" + std::to_string(save_counter) + ".csv");
            o << "i,j,lift,drag" << std::endl;
            for (std::size_t q = 0; q < considered_points.size() && 2 * q + 1 < per.size(); ++q)
                o << considered_points[q].coords[0] << "," << considered_points[q].coords[1] << "," << per[2 * q]
                  << "," << per[2 * q + 1] << std::endl;
            save_counter++;
        }
        lift_drag_per_iteration.emplace_back(lift, drag);
        return {lift, drag};
    }

    std::vector<ObstaclePoint<2>> considered_points;
    std::vector<int64_t> ij;
    std::vector<int32_t> masks;
    Pending pending[2];
    int next = 0;
    int64_t seq = 0;
    std::size_t iterations_between_save = 10;
    bool should_save = true;
    int save_counter = 0;
    std::vector<std::pair<double, double>> lift_drag_per_iteration;
};

} // namespace llalbm::analysis

#endif // LLALBM_HIP_FLOW_ANALYZER_HPP
