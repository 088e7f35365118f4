// Lattice and obstacle file readers and the velocity-field writer, in the
// reference's formats:
//  * lattice file (LatticeReader.hpp:300-522): "2" / "NR NC" / six counts
//    (InputNodeType order) / "i j type" for every non-fluid node;
//  * obstacle file (LatticeReader.hpp:532-575): count / "i j" per line;
//  * results (LatticeWriter.hpp:29-93): results/output-<k>u_x.txt and
//    u_y.txt, one space-separated row per lattice row.
// The readers produce a node-class plane (ColMajor, cls[i + NR*j]) -- the
// input of lbm_set_geometry -- in time linear in the file size.
#ifndef LLALBM_HIP_IO_HPP
#define LLALBM_HIP_IO_HPP

#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "aliases.hpp"

namespace llalbm::util::reader
{

/// Reads a lattice file into a class plane; dims receives {NR, NC}.
/// Throws std::runtime_error on a malformed file (the reference asserts).
inline std::vector<uint8_t> read_lattice_classes(const std::string &path, std::array<Index, 2> &dims)
{
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("the input file could not be opened: " + path);
    std::size_t file_dims = 0;
    in >> file_dims;
    if (file_dims != 2) throw std::runtime_error("file dimensions and template dimensions do not match");
    in >> dims[0] >> dims[1];
    if (!(dims[0] > 0 && dims[1] > 0)) throw std::runtime_error("extension of the dimensions must be positive");
    std::string line;
    std::getline(in, line);
    std::getline(in, line);
    std::istringstream counts_line(line);
    std::vector<std::size_t> counts;
    std::size_t v;
    while (counts_line >> v) counts.push_back(v);
    std::size_t total = 0;
    for (std::size_t c : counts) total += c;
    if (total != (std::size_t)(dims[0] * dims[1]))
        throw std::runtime_error("total number of elements and described number of nodes do not match");
    const std::size_t non_fluids = total - (counts.empty() ? 0 : counts[FLUID]);
    std::vector<uint8_t> cls((std::size_t)(dims[0] * dims[1]), FLUID);
    for (std::size_t q = 0; q < non_fluids; ++q) {
        Index i, j;
        std::size_t type;
        if (!(in >> i >> j >> type)) throw std::runtime_error("lattice file ended early");
        if (i < 0 || i >= dims[0] || j < 0 || j >= dims[1] || type > OBSTACLE)
            throw std::runtime_error("node outside the lattice or of unknown type");
        cls[(std::size_t)(i + dims[0] * j)] = (uint8_t)type;
    }
    return cls;
}

/// Reads an obstacle file and marks its nodes OBSTACLE in a class plane.
inline void read_obstacle_classes(const std::string &path, const std::array<Index, 2> &dims, std::vector<uint8_t> &cls)
{
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("the input file could not be opened: " + path);
    std::size_t n = 0;
    in >> n;
    for (std::size_t q = 0; q < n; ++q) {
        Index i, j;
        if (!(in >> i >> j)) throw std::runtime_error("obstacle file ended early");
        if (i < 0 || i >= dims[0] || j < 0 || j >= dims[1])
            throw std::runtime_error("obstacle node is not within the lattice bounds");
        cls[(std::size_t)(i + dims[0] * j)] = OBSTACLE;
    }
}

/// Obstacle file as a list of points (LatticeReader.hpp:532-575).
template <std::size_t dim>
void read_obstacle_file(const std::string &path, std::vector<ObstaclePoint<dim>> &obstacles,
                        const std::array<Index, dim> &dims)
{
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("the input file could not be opened: " + path);
    std::size_t n = 0;
    in >> n;
    for (std::size_t q = 0; q < n; ++q) {
        ObstaclePoint<dim> op;
        for (std::size_t k = 0; k < dim; ++k) {
            in >> op.coords[k];
            if (op.coords[k] < 0 || op.coords[k] >= dims[k])
                throw std::runtime_error("obstacle node is not within the lattice bounds");
        }
        obstacles.push_back(op);
    }
}

} // namespace llalbm::util::reader

namespace llalbm::util::writer
{

inline std::string output_dir_path = "./results/";

/// LatticeWriter.hpp:29-58: creates ./results/ or clears its .txt/.bin files.
inline void create_output_directory()
{
    std::filesystem::path path(output_dir_path);
    if (!std::filesystem::exists(path)) {
        std::filesystem::create_directory(path);
        return;
    }
    for (const auto &entry : std::filesystem::directory_iterator(path))
        if (entry.path().extension() == ".bin" || entry.path().extension() == ".txt")
            std::filesystem::remove(entry.path());
}

/// LatticeWriter.hpp:61-93: u_x and u_y as text rows, from a velocity
/// array in the ColMajor (NR, NC, 2) layout: (i, j, k) at i + NR*(j + NC*k).
inline void write_velocity_files(const double *u, Index nr, Index nc, std::size_t iteration_count)
{
    std::ofstream ux(output_dir_path + "/output-" + std::to_string(iteration_count) + "u_x.txt");
    std::ofstream uy(output_dir_path + "/output-" + std::to_string(iteration_count) + "u_y.txt");
    if (!ux.is_open() || !uy.is_open()) throw std::runtime_error("output file could not be opened");
    const double *vx = u, *vy = u + nr * nc;
    for (Index i = 0; i < nr; ++i) {
        for (Index j = 0; j < nc; ++j) {
            ux << vx[i + nr * j] << " ";
            uy << vy[i + nr * j] << " ";
        }
        ux << std::endl;
        uy << std::endl;
    }
}

/// LatticeWriter.hpp:61-93 on a velocity tensor.
template <std::size_t dim>
void write_lattice_file(const HostTensor<double, dim + 1> &global_u, std::size_t iteration_count)
{
    write_velocity_files(global_u.data(), global_u.dimension(0), global_u.dimension(1), iteration_count);
}

} // namespace llalbm::util::writer

#endif // LLALBM_HIP_IO_HPP
