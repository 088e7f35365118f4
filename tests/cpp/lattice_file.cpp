// A lattice read from a lattice file (the reference's 129 x 129 cavity
// fixture, test/assets/lid-lattice.txt) run through Lattice<HIPPolicy>, as
// the reference's test/unit/InletUpdates.cpp and LiddrivenTest.cpp build it.
//   lattice_file <lattice.txt> <steps> <save_step> <storage 0|1|2> <out prefix>
// Before the run it checks the getters and update_nodes on the host
// tensors (InletUpdates.cpp:33-44); the run saves u_x / u_y every save_step
// steps into ./results/ (save_step 0: no output); afterwards it writes
// <prefix>rho.bin, <prefix>u.bin (get_global_rho / get_global_u) and
// <prefix>pop.bin (get_populations) as raw doubles, and prints the storage
// scale of the context.
#include <llalbm_hip.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

namespace
{
template <typename T>
void dump(const std::string &path, const T &t)
{
    std::ofstream(path, std::ios::binary).write((const char *)t.data(), (std::streamsize)(t.size() * sizeof(double)));
}
} // namespace

int main(int argc, char **argv)
{
    using namespace llalbm::core;
    if (argc != 6) {
        std::fprintf(stderr, "usage: lattice_file lattice.txt steps save_step storage out_prefix\n");
        return 2;
    }
    const std::string path = argv[1], prefix = argv[5];
    const int steps = std::atoi(argv[2]);
    const int save_step = std::atoi(argv[3]);
    const int storage = std::atoi(argv[4]);

    using Configuration = LatticeConfiguration<2, collisions::HIPBGKCollisionPolicy<2>, boundaries::HIPBounceBackPolicy<2>,
                                               boundaries::HIPBounceBackPolicy<2>, boundaries::HIPZouHePolicy<2>,
                                               boundaries::HIPZouHePolicy<2>, initializers::HIPVelocityInitializer<2>,
                                               equilibrium::HIPDefaultEquilibrium<2>>;
    using Parallel = HIPPolicy<2, Configuration>;

    Lattice<Parallel> lid(path, 9, std::cout);
    lid.set_storage(storage);
    collisions::HIPBGKCollisionPolicy<2>::initialize(0.881);

    std::array<std::function<double(double, BoundaryPoint<2>)>, 2> lid_fns = {
        [](double t, BoundaryPoint<2>) { return 0.1 * (1.0 - std::exp(-t * 0.02)); },
        [](double, BoundaryPoint<2>) { return 0.0; }};
    std::array<std::function<double(double, BoundaryPoint<2>)>, 2> none = {
        [](double, BoundaryPoint<2>) { return 0.0; }, [](double, BoundaryPoint<2>) { return 0.0; }};
    initializers::HIPVelocityInitializer<2>::attach_update_functions(lid_fns, none);

    // before any run: rho = 1, u = 0, and update_nodes writes the values
    // of the update functions into the lattice's own velocity tensor
    if (lid.get_global_rho()(5, 7) != 1.0 || lid.get_global_u()(5, 7, 0) != 0.0) {
        std::fprintf(stderr, "getters before the run are not rho = 1, u = 0\n");
        return 1;
    }
    for (double t = 0; t < 2.0; t += 0.05) {
        const double val = lid_fns[0](t, BoundaryPoint<2>({0, 0}));
        initializers::HIPVelocityInitializer<2>::update_nodes(t, lid.get_global_u(), lid.get_global_rho());
        auto tensor = lid.get_global_u();
        for (Index j = 1; j < lid.get_lattice_dimensions()[1] - 1; ++j)
            if (tensor(0, j, 0) != val) {
                std::fprintf(stderr, "inlet velocity not updated at t = %g\n", t);
                return 1;
            }
    }
    std::printf("update_nodes checked\n");

    lid.perform_lbm(steps, 1, save_step > 0 ? (std::size_t)save_step : 1, save_step > 0);

    dump(prefix + "rho.bin", lid.get_global_rho());
    dump(prefix + "u.bin", lid.get_global_u());
    dump(prefix + "pop.bin", lid.get_populations());
    double scale = 0.0;
    lbm_get_scale(lid.context(), &scale);
    std::printf("scale %.17g inlets %zu\n", scale, lid.get_inlet_nodes().size());
    return 0;
}
