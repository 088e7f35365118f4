// Lattice<HIPPolicy<2, Configuration>>: the reference's Lattice
// (llalbm/src/core/lattices/Lattice.hpp:46-643) with its populations kept on
// the GPU.  Construction, perform_lbm, the getters and the flow-analyzer
// hook have the reference's names and meaning; each time step runs in the
// fused kernel behind the C ABI (lbm_step), and host copies are made only
// when a getter, a save step or an analysis step asks for them.
#ifndef LLALBM_HIP_LATTICE_HPP
#define LLALBM_HIP_LATTICE_HPP

#include <chrono>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/lbm_amd.h"
#include "aliases.hpp"
#include "flow_analyzer.hpp"
#include "io.hpp"
#include "policies.hpp"

namespace llalbm::core
{
using util::HostTensor;
using util::Index;

/// Error of a C-ABI call, with the library's message.
struct LbmError : std::runtime_error {
    int code;
    LbmError(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

inline void lbm_check(int code, const char *call)
{
    if (code != LBM_OK) throw LbmError(code, std::string(call) + ": " + lbm_last_error());
}

template <typename ParallelizationPolicy>
class Lattice
{
public:
    static constexpr std::size_t dim = 2;

    Lattice() = default;

    /// Lattice.hpp:173-205: an all-fluid lattice of the given extent.
    Lattice(const std::array<Index, 2> &dims, const std::size_t &q_, std::ostream &out_stream = std::cout)
        : q(q_), out(&out_stream)
    {
        std::vector<uint8_t> cls((std::size_t)(dims[0] * dims[1]), util::FLUID);
        reinit_from_classes(q_, dims, cls);
    }

    /// Lattice.hpp:207-239: a lattice read from a lattice file.
    Lattice(const std::string &path, const std::size_t &q_, std::ostream &out_stream = std::cout)
        : q(q_), out(&out_stream)
    {
        std::array<Index, 2> dims{};
        std::vector<uint8_t> cls = util::reader::read_lattice_classes(path, dims);
        reinit_from_classes(q_, dims, cls);
    }

    Lattice(const Lattice &) = delete;
    Lattice &operator=(const Lattice &) = delete;
    ~Lattice()
    {
        if (ctx) lbm_destroy(ctx);
    }

    /// Lattice.hpp:453-481 from node lists (the reference's reinit).
    void reinit(const std::size_t &q_, const std::array<Index, 2> &dims, const std::vector<Point<2>> &fluid,
                const std::vector<BoundaryPoint<2>> &boundary, const std::vector<BoundaryPoint<2>> &inlets,
                const std::vector<BoundaryPoint<2>> &outlets, const std::vector<ObstaclePoint<2>> &obstacles)
    {
        std::vector<uint8_t> cls((std::size_t)(dims[0] * dims[1]), util::SOLID);
        auto mark = [&](const auto &v, uint8_t k) {
            for (const auto &p : v) cls[(std::size_t)(p.coords[0] + dims[0] * p.coords[1])] = k;
        };
        mark(fluid, util::FLUID);
        mark(boundary, util::BOUNDARY);
        mark(inlets, util::INLET);
        mark(outlets, util::OUTLET);
        mark(obstacles, util::OBSTACLE);
        reinit_from_classes(q_, dims, cls);
    }

    /// Installs a node-class plane (ColMajor, cls[i + NR*j]).
    void reinit_from_classes(const std::size_t &q_, const std::array<Index, 2> &dims, const std::vector<uint8_t> &cls)
    {
        if (q_ != 9) throw std::invalid_argument("only D2Q9 (q = 9) is supported");
        q = q_;
        lattice_dimensions = dims;
        classes = cls;
        if (ctx) {
            lbm_destroy(ctx);
            ctx = nullptr;
            tuned = false;
        }
        build_lists();
        host_valid = false;
        ran = false;
        // Lattice.hpp:196-200: the initializer keeps copies of the lists
        ParallelizationPolicy::attach_nodes(inlet_nodes_coord, outlet_nodes_coord);
    }

    /// Lattice.hpp:288-438.  time / time_step steps; every save_step-th
    /// step writes results/output-<k>u_{x,y}.txt when should_save; an
    /// attached FlowAnalyzer runs at i > 0 && i % k == 0 on the
    /// post-collision populations.  Steps with no event and unchanged
    /// inlet/outlet values are launched together.  Save and analysis steps
    /// do not stall the GPU: their fields and reductions are downloaded
    /// asynchronously into page-locked double buffers, and a save is
    /// written to text while the GPU runs the steps queued after it.
    void perform_lbm(const double time, const double time_step = 1, const std::size_t save_step = 1,
                     const bool should_save = true)
    {
        if (!(time_step > 0) || !(time > 0) || !(time > time_step))
            throw std::invalid_argument("simulation time and time step must be positive, time > time step");
        if (should_save && save_step == 0) throw std::invalid_argument("save_step must be positive");
        ensure_context();
        if (should_save) util::writer::create_output_directory();
        if (flow_analyzer) flow_analyzer->initialize();
        const std::size_t n_steps = (std::size_t)(time / time_step);
        std::size_t saved_file = 0;
        std::vector<double> vals[2][3], last[2][3];
        // update_nodes (Lattice.hpp:351) of step i; the values of step 0 go
        // to the context before the initialisation, which picks the FP16
        // storage scale from them
        auto push_values = [&](std::size_t i, bool first, auto &&before_change) {
            for (int w = 0; w < 2; ++w) {
                const auto &nodes = w == 0 ? inlet_nodes_coord : outlet_nodes_coord;
                if (nodes.empty()) continue;
                ParallelizationPolicy::open_values((double)i, w, nodes, vals[w][0], vals[w][1], vals[w][2]);
                if (first || vals[w][0] != last[w][0] || vals[w][1] != last[w][1] || vals[w][2] != last[w][2]) {
                    before_change();
                    lbm_check(lbm_set_open_values(ctx, w, vals[w][0].data(), vals[w][1].data(), vals[w][2].data()),
                              "lbm_set_open_values");
                    for (int k = 0; k < 3; ++k) last[w][k] = vals[w][k];
                }
            }
        };
        if (n_steps > 0) push_values(0, true, [] {});
        lbm_check(lbm_init_equilibrium(ctx), "lbm_init_equilibrium");
        // the launch shape and occupancy, timed on the initial state (the
        // timing launches write only the spare buffer), before any
        // partially saturated bounce-back is attached (lbm_autotune refuses
        // those lattices)
        if (tune_steps > 0 && n_steps > 0 && !tuned) {
            lbm_check(lbm_autotune(ctx, tune_steps, nullptr, nullptr), "lbm_autotune");
            tuned = true;
        }
        ParallelizationPolicy::setup_obstacles(ctx, obstacle_nodes);
        saves.reset(lattice_dimensions[0], lattice_dimensions[1]);
        auto t0 = std::chrono::high_resolution_clock::now();
        std::int64_t pending = 0;
        // Queues the steps without events; the host then writes what has
        // arrived from earlier save and analysis steps while they run.
        auto flush = [&]() {
            if (pending) lbm_check(lbm_step(ctx, pending, step_flags()), "lbm_step");
            pending = 0;
            saves.write_all(ctx);
            if (flow_analyzer) flow_analyzer->collect();
        };
        for (std::size_t i = 0; i < n_steps; ++i) {
            if (i > 0) push_values(i, false, flush);
            const bool save = should_save && i % save_step == 0;
            const bool analyse = flow_analyzer && i > 0 && i % flow_analyzer->get_iterations_between_save() == 0;
            // the last step stores rho/u too: get_global_rho/u return them
            const bool final_step = i + 1 == n_steps;
            if (!save && !analyse && !final_step) {
                ++pending;
                continue;
            }
            if (pending) lbm_check(lbm_step(ctx, pending, step_flags()), "lbm_step");
            pending = 0;
            lbm_check(lbm_step(ctx, 1, step_flags() | ((save || final_step) ? LBM_STEP_MACRO : 0u)), "lbm_step");
            if (save) saves.queue(ctx, saved_file++);
            if (analyse) flow_analyzer->compute_on_device_async(ctx);
            // the GPU now runs this step (and those before it): write the
            // save before this one meanwhile
            saves.write_older(ctx);
        }
        flush();
        lbm_check(lbm_step(ctx, 0, LBM_STEP_CHECK), "lbm_step");
        lbm_check(lbm_synchronize(ctx), "lbm_synchronize");
        saves.write_all(ctx);
        if (flow_analyzer) flow_analyzer->collect();
        auto t1 = std::chrono::high_resolution_clock::now();
        last_elapsed = std::chrono::duration<double>(t1 - t0).count();
#ifdef TIMER
        std::cout << std::endl << "Time: " << last_elapsed << std::endl;
#endif
        if (flow_analyzer) flow_analyzer->save_global_results();
        host_valid = false;
        ran = n_steps > 0;
    }

    /// Lattice.hpp:492: the populations after the completed steps.
    HostTensor<double, 3> &get_populations()
    {
        ensure_context();
        populations = HostTensor<double, 3>(lattice_dimensions[0], lattice_dimensions[1], (Index)q);
        lbm_check(lbm_download_populations(ctx, populations.data()), "lbm_download_populations");
        return populations;
    }

    /// Lattice.hpp:499 / 506: rho and u of the last step of the last run
    /// (before any run: rho = 1, u = 0).  The tensors are the lattice's
    /// own: writes through the returned reference (update_nodes) persist
    /// until the next run.
    HostTensor<double, 3> &get_global_u()
    {
        if (!host_valid) download_macro();
        return global_u;
    }
    HostTensor<double, 2> &get_global_rho()
    {
        if (!host_valid) download_macro();
        return global_rho;
    }

    std::array<Index, 2> &get_lattice_dimensions() { return lattice_dimensions; }
    std::vector<Point<2>> &get_fluid_nodes() { return fluid_nodes; }
    std::vector<BoundaryPoint<2>> &get_boundary_nodes() { return boundary_coord; }
    std::vector<BoundaryPoint<2>> &get_inlet_nodes() { return inlet_nodes_coord; }
    std::vector<BoundaryPoint<2>> &get_outlet_nodes() { return outlet_nodes_coord; }
    std::vector<ObstaclePoint<2>> &get_obstacle_nodes() { return obstacle_nodes; }

    /// Lattice.hpp:556-627 (same sections and counts).
    void print_lattice_structure(std::ostream &o, const bool print_fluid = false) const
    {
        o << "===================================" << std::endl;
        o << "Lattice structure" << std::endl;
        o << "===================================" << std::endl;
        std::size_t n_nodes = 0;
        o << "-> List of obstacle nodes: " << std::endl;
        for (const auto &ob : obstacle_nodes) { o << ob.coords[0] << " " << ob.coords[1] << " dirs " << ob.directions << std::endl; n_nodes++; }
        o << "-> List of inlet nodes: " << std::endl;
        for (const auto &p : inlet_nodes_coord) { o << p.coords[0] << " " << p.coords[1] << " type " << p.type << std::endl; n_nodes++; }
        o << "-> List of outlet nodes: " << std::endl;
        for (const auto &p : outlet_nodes_coord) { o << p.coords[0] << " " << p.coords[1] << " type " << p.type << std::endl; n_nodes++; }
        o << "-> List of boundary nodes: " << std::endl;
        for (const auto &p : boundary_coord) { o << p.coords[0] << " " << p.coords[1] << " type " << p.type << std::endl; n_nodes++; }
        o << "===================================" << std::endl;
        o << "Cumulative number of non fluid: " << n_nodes << std::endl;
        o << "===================================" << std::endl;
        if (print_fluid) {
            o << "List of fluid nodes: " << std::endl;
            for (const auto &p : fluid_nodes) o << p.coords[0] << " " << p.coords[1] << std::endl;
        }
    }

    /// Lattice.hpp:245-278: weights b of the obstacles for partially
    /// saturated bounce-back, the fraction of obstacle nodes in each 5 x 5
    /// block.  The reference assigns them to copies of the points held in a
    /// std::map, so the lattice's obstacle list keeps its weights (b = 1
    /// unless set through get_obstacle_nodes()[k].set_b); this does the same,
    /// so results stay those of the reference.  The computed weights are
    /// returned, in obstacle list order.
    std::vector<double> compute_obstacle_weight()
    {
        std::map<std::pair<Index, Index>, std::size_t> count;
        for (const auto &o : obstacle_nodes) count[{o.coords[0] / 5, o.coords[1] / 5}]++;
        std::vector<double> w;
        w.reserve(obstacle_nodes.size());
        for (const auto &o : obstacle_nodes) w.push_back((double)count[{o.coords[0] / 5, o.coords[1] / 5}] / 25.0);
        return w;
    }

    /// Lattice.hpp:634-637.
    void attach_flow_analyzer(std::shared_ptr<analysis::FlowAnalyzer<2>> fa) { flow_analyzer = fa; }

    /// Launch-shape and occupancy tuning (no counterpart in the reference):
    /// with steps > 0, the next perform_lbm times every candidate over that
    /// many launches on its initial state and keeps the fastest (once per
    /// context; results are the same bits).  0 turns it off.
    void set_autotune(int steps = 5) { tune_steps = steps; }

    /// Storage of the populations (LBM_STORE_F64 is the reference-parity path).
    void set_storage(int32_t storage_)
    {
        storage = storage_;
        if (ctx) {
            lbm_destroy(ctx);
            ctx = nullptr;
            tuned = false;
        }
        host_valid = false;
        ran = false;
    }

    double elapsed_seconds() const { return last_elapsed; }
    lbm_ctx *context()
    {
        ensure_context();
        return ctx;
    }

private:
    static uint32_t step_flags()
    {
#ifdef LLALBM_DEBUG
        return LBM_STEP_CHECK;
#else
        return 0u;
#endif
    }

    void ensure_context()
    {
        if (lattice_dimensions[0] <= 0) throw std::logic_error("lattice has no geometry");
        if (ctx) {
            // relaxation parameters may have been set after construction
            lbm_desc d = ParallelizationPolicy::make_desc(lattice_dimensions[0], lattice_dimensions[1], storage);
            lbm_check(lbm_set_relaxation(ctx, d.tau, d.tau_even, d.tau_odd), "lbm_set_relaxation");
            return;
        }
        lbm_desc d = ParallelizationPolicy::make_desc(lattice_dimensions[0], lattice_dimensions[1], storage);
        lbm_check(lbm_create(&d, &ctx), "lbm_create");
        int r = lbm_set_geometry(ctx, classes.data());
        if (r != LBM_OK) {
            std::string msg = lbm_last_error();
            lbm_destroy(ctx);
            ctx = nullptr;
            tuned = false;
            throw LbmError(r, "lbm_set_geometry: " + msg);
        }
    }

    void download_macro()
    {
        global_rho = HostTensor<double, 2>(lattice_dimensions[0], lattice_dimensions[1]);
        global_u = HostTensor<double, 3>(lattice_dimensions[0], lattice_dimensions[1], (Index)2);
        if (ran) {
            ensure_context();
            lbm_check(lbm_download_macro(ctx, global_rho.data(), global_u.data()), "lbm_download_macro");
        } else {
            // Lattice.hpp:295-296 before the first step
            global_rho.setConstant(1.0);
            global_u.setConstant(0.0);
        }
        host_valid = true;
    }

    /// Page-locked double buffer of the u fields of save steps: a save
    /// queues the download of its step's fields (lbm_download_macro_async)
    /// and is written to text (LatticeWriter.hpp:61-93) by write_all, or
    /// when its buffer is needed again.
    class SaveQueue
    {
    public:
        SaveQueue() = default;
        SaveQueue(const SaveQueue &) = delete;
        SaveQueue &operator=(const SaveQueue &) = delete;
        ~SaveQueue() { release(); }

        void reset(Index nr_, Index nc_)
        {
            for (auto &sl : slot) sl.active = false;
            if (nr_ != nr || nc_ != nc) release();
            nr = nr_;
            nc = nc_;
        }
        void queue(lbm_ctx *ctx, std::size_t file_index)
        {
            Slot &sl = slot[next];
            if (sl.active) write(ctx, sl);
            if (!sl.u) {
                void *p = nullptr;
                lbm_check(lbm_host_alloc((int64_t)(2 * nr * nc * sizeof(double)), &p), "lbm_host_alloc");
                sl.u = (double *)p;
            }
            lbm_check(lbm_download_macro_async(ctx, nullptr, sl.u, &sl.ticket), "lbm_download_macro_async");
            sl.file = file_index;
            sl.seq = seq++;
            sl.active = true;
            next ^= 1;
        }
        /// Writes every queued save but the newest.
        void write_older(lbm_ctx *ctx)
        {
            if (slot[0].active && slot[1].active) write(ctx, slot[0].seq < slot[1].seq ? slot[0] : slot[1]);
        }
        /// Writes every queued save, oldest first.
        void write_all(lbm_ctx *ctx)
        {
            Slot *a = &slot[0], *b = &slot[1];
            if (b->active && (!a->active || b->seq < a->seq)) std::swap(a, b);
            if (a->active) write(ctx, *a);
            if (b->active) write(ctx, *b);
        }

    private:
        struct Slot {
            double *u = nullptr;
            int64_t ticket = 0, seq = 0;
            std::size_t file = 0;
            bool active = false;
        };
        void write(lbm_ctx *ctx, Slot &sl)
        {
            sl.active = false;
            lbm_check(lbm_wait(ctx, sl.ticket), "lbm_wait");
            util::writer::write_velocity_files(sl.u, nr, nc, sl.file);
        }
        void release()
        {
            for (auto &sl : slot) {
                lbm_host_free(sl.u);
                sl.u = nullptr;
                sl.active = false;
            }
        }
        Slot slot[2];
        int next = 0;
        int64_t seq = 0;
        Index nr = 0, nc = 0;
    };

    /// The reference's node lists: row-major order, btypes by position
    /// (NodeIdentification.hpp:45-110), obstacle direction masks and
    /// wet-obstacle elimination (:137-242).
    void build_lists()
    {
        fluid_nodes.clear();
        boundary_coord.clear();
        inlet_nodes_coord.clear();
        outlet_nodes_coord.clear();
        obstacle_nodes.clear();
        const Index nr = lattice_dimensions[0], nc = lattice_dimensions[1];
        static const int oi[9] = {0, 0, -1, 0, 1, -1, -1, 1, 1};
        static const int oj[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
        auto wet = [&](Index i, Index j) {
            uint8_t k = classes[(std::size_t)(i + nr * j)];
            return k == util::FLUID || k == util::INLET || k == util::OUTLET;
        };
        auto btype = [&](Index i, Index j) {
            if (i == 0) return j == 0 ? util::LEFT_TOP_C : j == nc - 1 ? util::RIGHT_TOP_C : util::TOP_WALL;
            if (i == nr - 1) return j == 0 ? util::LEFT_BOTTOM_C : j == nc - 1 ? util::RIGHT_BOTTOM_C : util::BOTTOM_WALL;
            if (j == 0) return util::LEFT_WALL;
            if (j == nc - 1) return util::RIGHT_WALL;
            return util::NONE;
        };
        for (Index i = 0; i < nr; ++i)
            for (Index j = 0; j < nc; ++j) {
                switch (classes[(std::size_t)(i + nr * j)]) {
                case util::FLUID: fluid_nodes.push_back(Point<2>({i, j})); break;
                case util::BOUNDARY: boundary_coord.push_back(BoundaryPoint<2>({i, j})); break;
                case util::INLET: {
                    BoundaryPoint<2> p({i, j});
                    p.type = btype(i, j);
                    inlet_nodes_coord.push_back(p);
                    break;
                }
                case util::OUTLET: {
                    BoundaryPoint<2> p({i, j});
                    p.type = btype(i, j);
                    outlet_nodes_coord.push_back(p);
                    break;
                }
                case util::OBSTACLE: {
                    ObstaclePoint<2> p({i, j});
                    for (int d = 1; d < 9; ++d) {
                        Index ii = i + oi[d], jj = j + oj[d];
                        if (ii >= 0 && ii < nr && jj >= 0 && jj < nc && wet(ii, jj)) p.directions.set(d);
                    }
                    if (p.directions.any()) obstacle_nodes.push_back(p);
                    break;
                }
                default: break;
                }
            }
    }

    std::size_t q = 0;
    std::ostream *out = &std::cout;
    std::array<Index, 2> lattice_dimensions{0, 0};
    std::vector<uint8_t> classes;
    std::vector<Point<2>> fluid_nodes;
    std::vector<BoundaryPoint<2>> boundary_coord, inlet_nodes_coord, outlet_nodes_coord;
    std::vector<ObstaclePoint<2>> obstacle_nodes;
    HostTensor<double, 3> populations, global_u;
    HostTensor<double, 2> global_rho;
    bool host_valid = false;
    bool ran = false;  // a run has completed since the geometry was set
    SaveQueue saves;
    int32_t storage = LBM_STORE_F64;
    int tune_steps = 0;  // set_autotune
    bool tuned = false;  // the context's shape and occupancy have been tuned
    lbm_ctx *ctx = nullptr;
    std::shared_ptr<analysis::FlowAnalyzer<2>> flow_analyzer;
    double last_elapsed = 0.0;
};

} // namespace llalbm::core

#endif // LLALBM_HIP_LATTICE_HPP
