// Lid-driven cavity written against the LLALBM API, as in the reference's
// examples/1_LidDrivenCavity129.cpp and test/unit/LidExecTest.cpp, with
// the HIP policies.  Writes the final populations (reference layout,
// binary) and the lift/drag history of an obstacle for the test suite.
#include <llalbm_hip.hpp>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

int main(int argc, char **argv)
{
    using namespace llalbm::core;
    using namespace llalbm::util::generation;
    const char *out_path = argc > 1 ? argv[1] : "populations.bin";
    const int steps = argc > 2 ? std::atoi(argv[2]) : 500;
    const bool tune = argc > 3 && std::string(argv[3]) == "tune";  // launch shape and occupancy tuned first

    using Configuration = LatticeConfiguration<2, collisions::HIPTRTCollisionPolicy<2>, boundaries::HIPBounceBackPolicy<2>,
                                               boundaries::HIPBounceBackPolicy<2>, boundaries::HIPZouHePolicy<2>,
                                               boundaries::HIPOpenBoundaryPolicy<2>, initializers::HIPVelocityInitializer<2>,
                                               equilibrium::HIPDefaultEquilibrium<2>>;
    using Parallelization = HIPPolicy<2, Configuration>;

    Lattice<Parallelization> lattice;
    ConstructionInfo<2> info;
    info.attach_domain_dimensions({65, 97});
    info.add_perimeter_nodes(NonFluidNodeType::BOUNDARY);
    info.add_nodes_interval({0, 1}, {0, 95}, NonFluidNodeType::INLET);
    info.add_obstacle_hyper_sphere({30, 40}, 6);
    build_lattice<2, Parallelization>(lattice, 9, info);

    collisions::HIPTRTCollisionPolicy<2>::initialize(0.58, 1.0, 1.0 / std::sqrt(3.0));
    collisions::HIPTRTCollisionPolicy<2>::enforce_magic_parameter();

    std::array<std::function<double(double, llalbm::util::BoundaryPoint<2>)>, 2> inlets = {
        [](double t, llalbm::util::BoundaryPoint<2> p) { return 0.1 * (1.0 - std::exp(-t * 0.01)); },
        [](double t, llalbm::util::BoundaryPoint<2> p) { return 0.0; }};
    std::array<std::function<double(double, llalbm::util::BoundaryPoint<2>)>, 2> outlets = {
        [](double t, llalbm::util::BoundaryPoint<2> p) { return 0.0; },
        [](double t, llalbm::util::BoundaryPoint<2> p) { return 0.0; }};
    initializers::HIPVelocityInitializer<2>::attach_update_functions(inlets, outlets);

    auto analysis = std::make_shared<FlowAnalyzer<2>>();
    analysis->add_point_vec(lattice.get_obstacle_nodes());
    analysis->set_should_save(false);
    analysis->set_iterations_between_save(100);
    lattice.attach_flow_analyzer(analysis);

    if (tune) lattice.set_autotune(3);
    lattice.perform_lbm(steps, 1, 100, false);

    auto &pop = lattice.get_populations();
    std::ofstream o(out_path, std::ios::binary);
    o.write((const char *)pop.data(), pop.size() * sizeof(double));
    for (const auto &ld : analysis->get_lift_drag_history()) std::printf("lift_drag %.17g %.17g\n", ld.first, ld.second);
    std::printf("obstacles %zu inlets %zu\n", lattice.get_obstacle_nodes().size(), lattice.get_inlet_nodes().size());
    return 0;
}
