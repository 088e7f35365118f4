// Host-only check of the LLALBM-compatible readers and writer: reads a
// lattice file and an obstacle file, writes their class planes as raw
// bytes, and writes a small velocity field through write_lattice_file.
//   io_check <lattice.txt> <classes.bin> <obstacles.txt> <nr> <nc> <obst_classes.bin> <results dir>
#include <cstdio>
#include <fstream>
#include <llalbm_hip/aliases.hpp>
#include <llalbm_hip/io.hpp>

using llalbm::util::Index;

int main(int argc, char **argv)
{
    if (argc != 8) {
        std::fprintf(stderr, "usage: io_check lattice classes.bin obstacles nr nc obst.bin results_dir\n");
        return 2;
    }
    std::array<Index, 2> dims{};
    auto cls = llalbm::util::reader::read_lattice_classes(argv[1], dims);
    std::ofstream(argv[2], std::ios::binary).write((const char *)cls.data(), (std::streamsize)cls.size());
    std::printf("dims %lld %lld\n", (long long)dims[0], (long long)dims[1]);

    const std::array<Index, 2> od{std::atoll(argv[4]), std::atoll(argv[5])};
    std::vector<uint8_t> oc((size_t)(od[0] * od[1]), 0);
    llalbm::util::reader::read_obstacle_classes(argv[3], od, oc);
    std::ofstream(argv[6], std::ios::binary).write((const char *)oc.data(), (std::streamsize)oc.size());
    std::vector<llalbm::util::ObstaclePoint<2>> pts;
    llalbm::util::reader::read_obstacle_file<2>(argv[3], pts, od);
    std::printf("obstacles %zu\n", pts.size());

    llalbm::util::writer::output_dir_path = argv[7];
    llalbm::util::writer::create_output_directory();
    llalbm::util::HostTensor<double, 3> u(3, 4, 2);
    for (Index i = 0; i < 3; ++i)
        for (Index j = 0; j < 4; ++j) {
            u(i, j, 0) = 0.25 * (double)(i * 4 + j);
            u(i, j, 1) = -(double)(i + j);
        }
    llalbm::util::writer::write_lattice_file<2>(u, 7);
    return 0;
}
