// The reference's test/unit/LidWithPSObstacle.cpp with the HIP policies: a
// 100 x 100 lid cavity (both lid velocity components ramped per side), a
// disc of radius 20, TRT 0.9 with magic 1/4, and partially saturated
// bounce-back as the obstacle policy with its tau from allowed_tau(0.02,
// 10) and the weights compute_obstacle_weight leaves (1, as the reference).
//   ps_obstacle <steps> <pop.bin>
#include <llalbm_hip.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>

int main(int argc, char **argv)
{
    using namespace llalbm::core;
    using namespace llalbm::util;
    if (argc != 3) {
        std::fprintf(stderr, "usage: ps_obstacle steps pop.bin\n");
        return 2;
    }
    const int steps = std::atoi(argv[1]);

    using Config = LatticeConfiguration<2, collisions::HIPTRTCollisionPolicy<2>, boundaries::HIPBounceBackPolicy<2>,
                                        boundaries::HIPPSBounceBackPolicy<2>, boundaries::HIPZouHePolicy<2>,
                                        boundaries::HIPZouHePolicy<2>, initializers::HIPVelocityInitializer<2>,
                                        equilibrium::HIPDefaultEquilibrium<2>>;
    using Parallel = HIPPolicy<2, Config>;
    Lattice<Parallel> lid;

    std::array<std::function<double(double, BoundaryPoint<2>)>, 2> velocity, outlets;
    velocity[1] = [](double time, BoundaryPoint<2> p) {
        if (p.coords[1] == 0) return 0.2 * (1.0 - std::exp(-((500 * 500 * time) / (2 * 1000 * 1000))));
        if (p.coords[1] == 99) return -0.2 * (1.0 - std::exp(-((500 * 500 * time) / (2 * 1000 * 1000))));
        return 0.0;
    };
    velocity[0] = [](double time, BoundaryPoint<2> p) {
        if (p.coords[0] == 0) return 0.2 * (1.0 - std::exp(-((500 * 500 * time) / (2 * 1000 * 1000))));
        if (p.coords[0] == 99) return -0.2 * (1.0 - std::exp(-((500 * 500 * time) / (2 * 1000 * 1000))));
        return 0.0;
    };
    initializers::HIPVelocityInitializer<2>::attach_update_functions(velocity, outlets);

    collisions::HIPTRTCollisionPolicy<2>::initialize(0.9, 0.01, 1. / std::sqrt(3.0));
    collisions::HIPTRTCollisionPolicy<2>::compute_magic_parameter();
    collisions::HIPTRTCollisionPolicy<2>::enforce_magic_parameter(1.0 / 4.0);

    generation::ConstructionInfo<2> info;
    info.attach_domain_dimensions({100, 100});
    info.add_perimeter_nodes(generation::NonFluidNodeType::BOUNDARY);
    info.add_nodes_interval({0, 1}, {0, 98}, generation::NonFluidNodeType::INLET);
    info.add_obstacle_hyper_sphere({50, 50}, 20);
    generation::build_lattice<2, Parallel>(lid, 9, info);

    boundaries::HIPPSBounceBackPolicy<2>::initialize(0.51, 0.01);
    boundaries::HIPPSBounceBackPolicy<2>::allowed_tau(0.02, 10);
    const auto weights = lid.compute_obstacle_weight();

    lid.perform_lbm(steps, 1, 10, false);

    auto &pop = lid.get_populations();
    std::ofstream(argv[2], std::ios::binary).write((const char *)pop.data(), (std::streamsize)(pop.size() * sizeof(double)));
    double wmin = 1.0;
    for (double w : weights) wmin = std::min(wmin, w);
    std::printf("obstacles %zu tau_ps %.17g min_block_weight %.17g b0 %.17g\n", lid.get_obstacle_nodes().size(),
                boundaries::HIPPSBounceBackPolicy<2>::tau, wmin, lid.get_obstacle_nodes()[0].get_b());
    return 0;
}
