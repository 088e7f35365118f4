// HIP-tagged policy classes and the HIPPolicy parallelization policy.
//
// The reference selects its numerics at compile time: a
// LatticeConfiguration bundles seven policy types
// (llalbm/src/core/lattices/LatticeConfiguration.hpp:32-60), all carrying the
// same parallelization tag, and a parallelization policy such as
// OMPPolicy<2, Configuration> (parallelization/OMPPolicy.hpp:36-164)
// dispatches Lattice::perform_lbm's calls to them.  This header adds the
// HIP tag, the IsGloballyHIP concept (PolicyTypes.hpp:72-125 pattern) and
// HIP-tagged policy classes with the same static parameter setters, so a
// reference configuration switches to the GPU by renaming its policies.
// The numerics themselves run in the C ABI (lbm_amd.h); these classes hold
// parameters and translate them into an lbm_desc.
#ifndef LLALBM_HIP_POLICIES_HPP
#define LLALBM_HIP_POLICIES_HPP

#include <array>
#include <concepts>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/lbm_amd.h"
#include "aliases.hpp"

namespace llalbm::core
{
using util::BoundaryPoint;
using util::ObstaclePoint;
using util::Point;

// ----- kind and parallelization tags (PolicyTypes.hpp:36-45) -----
class CollisionPolicyTag {};
class BoundaryPolicyTag {};
class InitializationPolicyTag {};
class HIPTag {};

template <typename T>
concept IsCollisionPolicy = std::is_base_of_v<CollisionPolicyTag, T>;
template <typename T>
concept IsBoundaryPolicy = std::is_base_of_v<BoundaryPolicyTag, T>;
template <typename T>
concept IsInitializationPolicy = std::is_base_of_v<InitializationPolicyTag, T>;
template <typename T>
concept IsHIP = std::is_base_of_v<HIPTag, T>;

// ----- LatticeConfiguration (LatticeConfiguration.hpp:32-60) -----
template <std::size_t dim, typename CollisionPolicy, typename WallPolicy, typename ObstaclePolicy,
          typename InletPolicy, typename OutletPolicy, typename InitializationPolicy, typename EquilibriumPolicy>
    requires IsCollisionPolicy<CollisionPolicy> && IsBoundaryPolicy<WallPolicy> && IsBoundaryPolicy<ObstaclePolicy> &&
             IsBoundaryPolicy<InletPolicy> && IsBoundaryPolicy<OutletPolicy> &&
             IsInitializationPolicy<InitializationPolicy>
class LatticeConfiguration
{
public:
    static constexpr std::size_t dimensions = dim;
    using collision_policy_t = CollisionPolicy;
    using wall_policy_t = WallPolicy;
    using obstacle_policy_t = ObstaclePolicy;
    using inlet_policy_t = InletPolicy;
    using outlet_policy_t = OutletPolicy;
    using initialization_policy_t = InitializationPolicy;
    using equilibrium_policy_t = EquilibriumPolicy;
};

template <typename Configuration>
concept IsGloballyHIP = std::conjunction_v<std::is_base_of<HIPTag, typename Configuration::collision_policy_t>,
                                           std::is_base_of<HIPTag, typename Configuration::wall_policy_t>,
                                           std::is_base_of<HIPTag, typename Configuration::obstacle_policy_t>,
                                           std::is_base_of<HIPTag, typename Configuration::inlet_policy_t>,
                                           std::is_base_of<HIPTag, typename Configuration::outlet_policy_t>,
                                           std::is_base_of<HIPTag, typename Configuration::initialization_policy_t>,
                                           std::is_base_of<HIPTag, typename Configuration::equilibrium_policy_t>>;

namespace collisions
{
template <std::size_t dim>
class HIPBGKCollisionPolicy;

/// BGK collision on the GPU; parameters as BGKCollisionPolicy<2>
/// (BGKCollision.hpp:67-107).
template <>
class HIPBGKCollisionPolicy<2> : public CollisionPolicyTag, public HIPTag
{
public:
    static inline double tau = 0.0;
    static constexpr int32_t kind = LBM_BGK;

    static void initialize(const double &tau_, const double &max_speed = 0.2)
    {
        tau = tau_;
        std::cout << "Inizialization done:" << std::endl;
        std::cout << " Tau: " << tau << std::endl;
        std::cout << " Max Speed: " << max_speed << std::endl;
        if (tau <= 0.55) std::cout << " The simulation might have stability issues." << std::endl;
    }

    static void allowed_tau(const double &max_speed, const double &reynolds, const double &cs2 = 1.0 / 3.0)
    {
        tau = 0.5 + max_speed / (cs2 * reynolds);
    }

    static void fill(lbm_desc &d) { d.collision = LBM_BGK; d.tau = tau; }
};

template <std::size_t dim>
class HIPTRTCollisionPolicy;

/// TRT collision on the GPU; parameters as TRTCollisionPolicy<2>
/// (TRTCollision.hpp:69-134).
template <>
class HIPTRTCollisionPolicy<2> : public CollisionPolicyTag, public HIPTag
{
public:
    static inline double tau_even = 0.0;
    static inline double tau_odd = 0.0;
    static inline double sound_velocity = 0.0;
    static constexpr int32_t kind = LBM_TRT;

    static void initialize(const double &tau_even_, const double &tau_odd_, const double &sound_velocity_)
    {
        tau_even = tau_even_;
        tau_odd = tau_odd_;
        sound_velocity = sound_velocity_;
    }
    static double compute_kinematic_viscosity() { return (tau_even - 0.5) * sound_velocity * sound_velocity; }
    static double compute_magic_parameter() { return (tau_even - 0.5) * (tau_odd - 0.5); }
    static void enforce_magic_parameter(const double &magic_parameter = 1.0 / 4.0)
    {
        tau_odd = (magic_parameter) * (1.0 / (tau_even - 0.5)) + 0.5;
    }

    static void fill(lbm_desc &d) { d.collision = LBM_TRT; d.tau_even = tau_even; d.tau_odd = tau_odd; }
};
} // namespace collisions

namespace boundaries
{
template <std::size_t dim>
class HIPBounceBackPolicy;
/// Half-way bounce-back on walls and wet obstacles (BounceBackPolicy.hpp:68-187).
template <>
class HIPBounceBackPolicy<2> : public BoundaryPolicyTag, public HIPTag
{
public:
    static constexpr bool partially_saturated = false;
};

template <std::size_t dim>
class HIPPSBounceBackPolicy;
/// Partially saturated bounce-back as the obstacle policy
/// (PSBounceBackPolicy<2>, PSBounceBackPolicy.hpp:60-297): walls bounce
/// back as usual, each wet obstacle mixes bounce-back and the neighbour's
/// equilibrium with its weight b (ObstaclePoint<2>::b).  The policy's own
/// relaxation time is set by initialize / allowed_tau, as in the reference.
/// FP64 storage only.
template <>
class HIPPSBounceBackPolicy<2> : public BoundaryPolicyTag, public HIPTag
{
public:
    static inline double tau = 0.0;
    static constexpr bool partially_saturated = true;

    /// PSBounceBackPolicy.hpp:66-79.
    static void initialize(const double &tau_, const double &max_speed = 0.2)
    {
        tau = tau_;
        std::cout << "Inizialization done:" << std::endl;
        std::cout << " Tau: " << tau << std::endl;
        std::cout << " Max Speed: " << max_speed << std::endl;
        if (tau <= 0.55) std::cout << " The simulation might have stability issues." << std::endl;
        std::cout << " Allowed Reynolds Number: " << max_speed * 3.0 / (tau - 0.5) << std::endl;
    }

    /// PSBounceBackPolicy.hpp:89-98.
    static void allowed_tau(const double &max_speed, const double &reynolds, const double &speed_of_sound_squared = 1.0 / 3.0)
    {
        tau = 0.5 + max_speed / (speed_of_sound_squared * reynolds);
    }
};

template <std::size_t dim>
class HIPZouHePolicy;
/// Zou-He velocity boundary on an inlet/outlet list (ZouHePolicy.hpp:58-123).
template <>
class HIPZouHePolicy<2> : public BoundaryPolicyTag, public HIPTag
{
public:
    static constexpr int32_t open_policy = LBM_ZOUHE_VELOCITY;
};

template <std::size_t dim>
class HIPZouHePressurePolicy;
/// Zou-He pressure boundary (no reference counterpart): density prescribed
/// by the initializer's density functions.
template <>
class HIPZouHePressurePolicy<2> : public BoundaryPolicyTag, public HIPTag
{
public:
    static constexpr int32_t open_policy = LBM_ZOUHE_PRESSURE;
};

template <std::size_t dim>
class HIPOpenBoundaryPolicy;
/// No-op open boundary (OpenBoundaryPolicy.hpp:58-61).
template <>
class HIPOpenBoundaryPolicy<2> : public BoundaryPolicyTag, public HIPTag
{
public:
    static constexpr int32_t open_policy = LBM_OPEN_NOOP;
};
} // namespace boundaries

namespace initializers
{
template <std::size_t dim>
class HIPVelocityInitializer;

/// Prescribed inlet/outlet values, evaluated on the host once per step
/// (VelocityInitializer<2>, InletOutletInitializer.hpp:34-156).  The
/// density functions feed HIPZouHePressurePolicy lists.
template <>
class HIPVelocityInitializer<2> : public InitializationPolicyTag, public HIPTag
{
public:
    using Fn = std::function<double(double, BoundaryPoint<2>)>;
    static inline std::vector<BoundaryPoint<2>> inlet_nodes{};
    static inline std::vector<BoundaryPoint<2>> outlet_nodes{};
    static inline std::array<Fn, 2> inlet_update_function{};
    static inline std::array<Fn, 2> outlet_update_function{};
    static inline Fn inlet_density_function{};
    static inline Fn outlet_density_function{};

    /// InletOutletInitializer.hpp:60-67: copies of the lattice's lists.
    static void attach_nodes(std::vector<BoundaryPoint<2>> &inlet_nodes_, std::vector<BoundaryPoint<2>> &outlet_nodes_)
    {
        inlet_nodes = inlet_nodes_;
        outlet_nodes = outlet_nodes_;
    }

    static void attach_update_functions(std::array<Fn, 2> &inlet_fns, std::array<Fn, 2> &outlet_fns)
    {
        inlet_update_function = inlet_fns;
        outlet_update_function = outlet_fns;
    }

    /// InletOutletInitializer.hpp:90-129: the velocity tensor takes the
    /// update functions' values at time t on the attached inlet and outlet
    /// nodes.  The density tensor is not written (the reference receives
    /// it by value).
    static void update_nodes(const double &t, util::HostTensor<double, 3> &velocity_tensor,
                             const util::HostTensor<double, 2> &density_tensor)
    {
        (void)density_tensor;
        for (int w = 0; w < 2; ++w) {
            const auto &nodes = w == 0 ? inlet_nodes : outlet_nodes;
            const auto &fns = w == 0 ? inlet_update_function : outlet_update_function;
            for (const auto &p : nodes)
                for (int k = 0; k < 2; ++k)
                    if (fns[k]) velocity_tensor(p.coords[0], p.coords[1], (util::Index)k) = fns[k](t, p);
        }
    }
    static void attach_density_functions(const Fn &inlet_rho, const Fn &outlet_rho)
    {
        inlet_density_function = inlet_rho;
        outlet_density_function = outlet_rho;
    }
};
} // namespace initializers

namespace equilibrium
{
template <std::size_t dim>
class HIPDefaultEquilibrium;
/// Second-order D2Q9 equilibrium (DefaultEquilibrium.hpp:54-155).
template <>
class HIPDefaultEquilibrium<2> : public HIPTag
{
};
} // namespace equilibrium

template <std::size_t dim, typename Configuration>
class HIPPolicy;

/// GPU parallelization policy.  Where OMPPolicy forwards each of
/// perform_lbm's calls to a policy class operating on host tensors, this
/// policy describes the whole configuration to the C ABI once
/// (make_desc) and evaluates the inlet/outlet functions per step
/// (open_values); Lattice<HIPPolicy<2, C>> then runs the fused step.
///
/// It also has the reference policy's statics (OMPPolicy.hpp:50-163), so
/// a loop written like the reference's perform_lbm (Lattice.hpp:338-410)
/// runs on the GPU once a context is bound (bind).  The populations stay
/// on the device, so the tensors the statics receive are not read:
///  * update_nodes evaluates the update functions into the host velocity
///    tensor (as the reference does) and into the context;
///  * the first calc_equilibrium after bind is the run's prologue
///    (Lattice.hpp:322-328) and initialises the device state;
///  * collide launches the whole fused time step (lbm_step);
///  * update_macro, the later calc_equilibrium calls, stream, the open
///    collide/stream calls and the four boundary updates are work the
///    fused step has done, and do nothing.
template <typename Configuration>
    requires IsGloballyHIP<Configuration>
class HIPPolicy<2, Configuration> : public HIPTag
{
public:
    static constexpr std::size_t dims = Configuration::dimensions;
    using CollisionPolicy = typename Configuration::collision_policy_t;
    using InletPolicy = typename Configuration::inlet_policy_t;
    using OutletPolicy = typename Configuration::outlet_policy_t;
    using ObstaclePolicy = typename Configuration::obstacle_policy_t;
    using InitializationPolicy = typename Configuration::initialization_policy_t;
    using Tensor3 = util::HostTensor<double, 3>;
    using Tensor2 = util::HostTensor<double, 2>;

    /// The context the statics act on, its obstacle list, and whether its
    /// run has started.
    static inline lbm_ctx *context = nullptr;
    static inline const std::vector<ObstaclePoint<2>> *obstacles = nullptr;
    static inline bool started = false;
    /// lbm_step flags of the launches collide makes.
    static inline uint32_t step_flags = 0;

    static void bind(lbm_ctx *ctx, const std::vector<ObstaclePoint<2>> *obstacle_list = nullptr)
    {
        context = ctx;
        obstacles = obstacle_list;
        started = false;
    }

    /// The obstacle policy's set-up at the start of a run: partially
    /// saturated bounce-back hands its tau and the obstacles' weights (list
    /// order) to the context; plain bounce-back needs nothing.
    static void setup_obstacles(lbm_ctx *ctx, const std::vector<ObstaclePoint<2>> &list)
    {
        if constexpr (requires { ObstaclePolicy::partially_saturated; }) {
            if constexpr (ObstaclePolicy::partially_saturated) {
                std::vector<double> b(list.size());
                for (std::size_t k = 0; k < list.size(); ++k) b[k] = list[k].get_b();
                check(lbm_set_partial_bounce_back(ctx, ObstaclePolicy::tau, b.data()),
                      "lbm_set_partial_bounce_back");
            }
        }
    }

    // ----- update of the boundaries (OMPPolicy.hpp:50-72): in the fused step
    static void update_inlet_boundaries(Tensor3 &, std::vector<BoundaryPoint<2>> &, const Tensor2 &, const Tensor3 &) {}
    static void update_outlet_boundaries(Tensor3 &, std::vector<BoundaryPoint<2>> &, const Tensor2 &, const Tensor3 &) {}
    static void update_domain_boundaries(Tensor3 &, std::vector<BoundaryPoint<2>> &, const Tensor2 &, const Tensor3 &) {}
    static void update_obstacle_boundaries(Tensor3 &, std::vector<ObstaclePoint<2>> &, const Tensor2 &, const Tensor3 &) {}

    // ----- collision and streaming (OMPPolicy.hpp:82-115)
    /// One whole time step on the bound context.
    static void collide(const Tensor3 &, Tensor3 &, Tensor3 &, const std::vector<Point<2>> &, Tensor2 &, Tensor3 &,
                        const double)
    {
        require_context();
        check(lbm_step(context, 1, step_flags), "lbm_step");
    }
    static void stream(Tensor3 &, Tensor3 &, const std::vector<Point<2>> &) {}
    static void collide_open_inlet_boundary(const Tensor3 &, Tensor3 &, Tensor3 &, const std::vector<BoundaryPoint<2>> &,
                                            Tensor2 &, Tensor3 &, const double)
    {
    }
    static void collide_open_outlet_boundary(const Tensor3 &, Tensor3 &, Tensor3 &, const std::vector<BoundaryPoint<2>> &,
                                             Tensor2 &, Tensor3 &, const double)
    {
    }
    static void stream_open_inlet_boundary(Tensor3 &, Tensor3 &, const std::vector<BoundaryPoint<2>> &) {}
    static void stream_open_outlet_boundary(Tensor3 &, Tensor3 &, const std::vector<BoundaryPoint<2>> &) {}

    // ----- initialisation and external updates (OMPPolicy.hpp:127-143)
    static void attach_nodes(std::vector<BoundaryPoint<2>> &inlet_nodes_, std::vector<BoundaryPoint<2>> &outlet_nodes_)
    {
        InitializationPolicy::attach_nodes(inlet_nodes_, outlet_nodes_);
    }
    static void update_macro(const Tensor3 &, const std::vector<Point<2>> &, Tensor2 &, Tensor3 &) {}
    /// Values of step t into the velocity tensor and the bound context.
    static void update_nodes(const double &t, Tensor3 &velocity_tensor, const Tensor2 &density_tensor)
    {
        InitializationPolicy::update_nodes(t, velocity_tensor, density_tensor);
        if (!context) return;
        std::vector<double> ux, uy, rho;
        for (int w = 0; w < 2; ++w) {
            const auto &nodes = w == 0 ? InitializationPolicy::inlet_nodes : InitializationPolicy::outlet_nodes;
            if (nodes.empty()) continue;
            open_values(t, w, nodes, ux, uy, rho);
            check(lbm_set_open_values(context, w, ux.data(), uy.data(), rho.data()), "lbm_set_open_values");
        }
    }

    // ----- equilibrium (OMPPolicy.hpp:153-163)
    static void calc_equilibrium(std::vector<Point<2>> &, Tensor3 &, Tensor3 &, Tensor2 &) { prologue(); }
    static void calc_equilibrium(std::vector<BoundaryPoint<2>> &, Tensor3 &, Tensor3 &, Tensor2 &) { prologue(); }

    static lbm_desc make_desc(util::Index nr, util::Index nc, int32_t storage = LBM_STORE_F64, int32_t device = 0)
    {
        lbm_desc d{};
        d.n_rows = nr;
        d.n_cols = nc;
        d.storage = storage;
        d.tau = d.tau_even = d.tau_odd = 1.0;
        CollisionPolicy::fill(d);
        d.inlet_policy = InletPolicy::open_policy;
        d.outlet_policy = OutletPolicy::open_policy;
        d.device = device;
        return d;
    }

    /// Values of the update functions at time t for the inlet (which = 0)
    /// or outlet (1) list (InletOutletInitializer.hpp:90-129).
    static void open_values(double t, int which, const std::vector<BoundaryPoint<2>> &nodes, std::vector<double> &ux,
                            std::vector<double> &uy, std::vector<double> &rho)
    {
        const auto &fns = which == 0 ? InitializationPolicy::inlet_update_function
                                     : InitializationPolicy::outlet_update_function;
        const auto &rfn = which == 0 ? InitializationPolicy::inlet_density_function
                                     : InitializationPolicy::outlet_density_function;
        ux.resize(nodes.size());
        uy.resize(nodes.size());
        rho.resize(nodes.size());
        for (std::size_t q = 0; q < nodes.size(); ++q) {
            ux[q] = fns[0] ? fns[0](t, nodes[q]) : 0.0;
            uy[q] = fns[1] ? fns[1](t, nodes[q]) : 0.0;
            rho[q] = rfn ? rfn(t, nodes[q]) : 1.0;
        }
    }

private:
    static void require_context()
    {
        if (!context) throw std::logic_error("HIPPolicy: no context bound (HIPPolicy::bind)");
    }
    static void check(int code, const char *call)
    {
        if (code != LBM_OK) throw std::runtime_error(std::string(call) + ": " + lbm_last_error());
    }
    static void prologue()
    {
        require_context();
        if (started) return;
        check(lbm_init_equilibrium(context), "lbm_init_equilibrium");
        if (obstacles) setup_obstacles(context, *obstacles);
        started = true;
    }
};

} // namespace llalbm::core

#endif // LLALBM_HIP_POLICIES_HPP
