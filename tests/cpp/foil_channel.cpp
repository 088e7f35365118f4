// Flow past the NACA aerofoil of the reference's examples/NACA-foil.txt in a
// 100 x 400 channel, written like examples/8_1_NACAProfileFlight.cpp with
// the HIP policies: perimeter walls, Zou-He velocity inlet on column 0 with
// a ramped speed, no-op outlet on column 399, the aerofoil read by
// ConstructionInfo::read_obstacle_from_file, TRT, and a FlowAnalyzer on
// every obstacle node that saves its per-point contributions.
//   foil_channel <NACA-foil.txt> <steps> <save_step> <analysis_every> <pop.bin>
// Output: ./results/output-<k>u_{x,y}.txt (save_step > 0),
// ./flow_analysis/contribution_<k>.csv and total_lift_drag.csv, and the
// final populations as raw doubles.
#include <llalbm_hip.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

int main(int argc, char **argv)
{
    using namespace llalbm::core;
    using namespace llalbm::util::generation;
    if (argc != 6) {
        std::fprintf(stderr, "usage: foil_channel NACA-foil.txt steps save_step analysis_every pop.bin\n");
        return 2;
    }
    const int steps = std::atoi(argv[2]);
    const int save_step = std::atoi(argv[3]);
    const int every = std::atoi(argv[4]);

    using Configuration = LatticeConfiguration<2, collisions::HIPTRTCollisionPolicy<2>, boundaries::HIPBounceBackPolicy<2>,
                                               boundaries::HIPBounceBackPolicy<2>, boundaries::HIPZouHePolicy<2>,
                                               boundaries::HIPOpenBoundaryPolicy<2>, initializers::HIPVelocityInitializer<2>,
                                               equilibrium::HIPDefaultEquilibrium<2>>;
    using Parallelization = HIPPolicy<2, Configuration>;

    Lattice<Parallelization> lattice;
    ConstructionInfo<2> info;
    info.attach_domain_dimensions({100, 400});
    info.add_perimeter_nodes(NonFluidNodeType::BOUNDARY);
    info.add_nodes_interval({1, 0}, {98, 0}, NonFluidNodeType::INLET);
    info.add_nodes_interval({1, 399}, {98, 399}, NonFluidNodeType::OUTLET);
    info.read_obstacle_from_file(argv[1]);
    build_lattice<2, Parallelization>(lattice, 9, info);

    collisions::HIPTRTCollisionPolicy<2>::initialize(0.9, 1.2, 1.0 / std::sqrt(3.0));
    std::array<std::function<double(double, BoundaryPoint<2>)>, 2> inlets = {
        [](double t, BoundaryPoint<2>) { return 0.15 * (1 - std::exp(-t * 0.001)); },
        [](double, BoundaryPoint<2>) { return 0.0; }};
    std::array<std::function<double(double, BoundaryPoint<2>)>, 2> outlets = {
        [](double, BoundaryPoint<2>) { return 0.0; }, [](double, BoundaryPoint<2>) { return 0.0; }};
    initializers::HIPVelocityInitializer<2>::attach_update_functions(inlets, outlets);

    auto analysis = std::make_shared<FlowAnalyzer<2>>();
    analysis->add_point_vec(lattice.get_obstacle_nodes());
    analysis->set_should_save(true);
    analysis->set_iterations_between_save((std::size_t)every);
    lattice.attach_flow_analyzer(analysis);

    lattice.perform_lbm(steps, 1, save_step > 0 ? (std::size_t)save_step : 1, save_step > 0);

    auto &pop = lattice.get_populations();
    std::ofstream(argv[5], std::ios::binary).write((const char *)pop.data(), (std::streamsize)(pop.size() * sizeof(double)));
    std::printf("obstacles %zu inlets %zu outlets %zu analyses %zu\n", lattice.get_obstacle_nodes().size(),
                lattice.get_inlet_nodes().size(), lattice.get_outlet_nodes().size(),
                analysis->get_lift_drag_history().size());
    return 0;
}
