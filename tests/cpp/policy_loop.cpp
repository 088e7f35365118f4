// The reference's perform_lbm loop body (Lattice.hpp:322-410), written
// out against the parallelization-policy statics the way Lattice calls
// them, with HIPPolicy bound to the lattice's context: a channel with a
// cylinder, TRT, Zou-He velocity inlet (time-dependent) and no-op outlet.
//   policy_loop <steps> <pop.bin>
// Writes the populations after the steps (raw doubles) and prints the
// inlet velocity the host tensor holds after the last update_nodes.
#include <llalbm_hip.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>

int main(int argc, char **argv)
{
    using namespace llalbm::core;
    using namespace llalbm::util::generation;
    if (argc != 3) {
        std::fprintf(stderr, "usage: policy_loop steps pop.bin\n");
        return 2;
    }
    const int steps = std::atoi(argv[1]);

    using Configuration = LatticeConfiguration<2, collisions::HIPTRTCollisionPolicy<2>, boundaries::HIPBounceBackPolicy<2>,
                                               boundaries::HIPBounceBackPolicy<2>, boundaries::HIPZouHePolicy<2>,
                                               boundaries::HIPOpenBoundaryPolicy<2>, initializers::HIPVelocityInitializer<2>,
                                               equilibrium::HIPDefaultEquilibrium<2>>;
    using P = HIPPolicy<2, Configuration>;

    Lattice<P> lattice;
    ConstructionInfo<2> info;
    info.attach_domain_dimensions({48, 120});
    info.add_perimeter_nodes(NonFluidNodeType::BOUNDARY);
    info.add_nodes_interval({1, 0}, {46, 0}, NonFluidNodeType::INLET);
    info.add_nodes_interval({1, 119}, {46, 119}, NonFluidNodeType::OUTLET);
    info.add_obstacle_hyper_sphere({22, 30}, 6);
    build_lattice<2, P>(lattice, 9, info);

    collisions::HIPTRTCollisionPolicy<2>::initialize(0.6, 1.0, 1.0 / std::sqrt(3.0));
    collisions::HIPTRTCollisionPolicy<2>::enforce_magic_parameter(3.0 / 16.0);
    std::array<std::function<double(double, BoundaryPoint<2>)>, 2> inlets = {
        [](double t, BoundaryPoint<2> p) {
            const double y = (double)p.coords[0] - 0.5, h = 46.0;
            return 4.0 * 0.05 * y * (h - y) / (h * h) * (1.0 - std::exp(-t * 0.01));
        },
        [](double, BoundaryPoint<2>) { return 0.0; }};
    std::array<std::function<double(double, BoundaryPoint<2>)>, 2> outlets = {
        [](double, BoundaryPoint<2>) { return 0.0; }, [](double, BoundaryPoint<2>) { return 0.0; }};
    initializers::HIPVelocityInitializer<2>::attach_update_functions(inlets, outlets);

    // the tensors the reference's loop hands to the policy
    const auto dims = lattice.get_lattice_dimensions();
    llalbm::util::HostTensor<double, 3> populations(dims[0], dims[1], (Index)9), equilibrium(dims[0], dims[1], (Index)9),
        after(dims[0], dims[1], (Index)9);
    auto &rho = lattice.get_global_rho();
    auto &u = lattice.get_global_u();
    auto &fluid = lattice.get_fluid_nodes();
    auto &walls = lattice.get_boundary_nodes();
    auto &in = lattice.get_inlet_nodes();
    auto &out = lattice.get_outlet_nodes();
    auto &obst = lattice.get_obstacle_nodes();

    P::bind(lattice.context(), &obst);
    rho.setConstant(1);
    u.setConstant(0);
    P::calc_equilibrium(fluid, populations, u, rho);
    P::calc_equilibrium(in, populations, u, rho);
    P::calc_equilibrium(out, populations, u, rho);
    for (int i = 0; i < steps; ++i) {
        P::update_nodes(i, u, rho);
        P::update_macro(populations, fluid, rho, u);
        P::calc_equilibrium(fluid, equilibrium, u, rho);
        P::calc_equilibrium(in, equilibrium, u, rho);
        P::calc_equilibrium(out, equilibrium, u, rho);
        P::collide(populations, equilibrium, after, fluid, rho, u, 1.0);
        P::collide_open_inlet_boundary(populations, equilibrium, after, in, rho, u, 1.0);
        P::collide_open_outlet_boundary(populations, equilibrium, after, out, rho, u, 1.0);
        P::stream(populations, after, fluid);
        P::stream_open_inlet_boundary(populations, after, in);
        P::stream_open_outlet_boundary(populations, after, out);
        P::update_domain_boundaries(populations, walls, rho, u);
        P::update_obstacle_boundaries(populations, obst, rho, u);
        P::update_inlet_boundaries(populations, in, rho, u);
        P::update_outlet_boundaries(populations, out, rho, u);
    }
    auto &pop = lattice.get_populations();
    std::ofstream(argv[2], std::ios::binary).write((const char *)pop.data(), (std::streamsize)(pop.size() * sizeof(double)));
    std::printf("inlet u(10, 0) %.17g\n", u(10, 0, 0));
    return 0;
}
