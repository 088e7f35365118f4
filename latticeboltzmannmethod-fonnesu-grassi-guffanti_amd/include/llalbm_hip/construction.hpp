// In-code lattice construction with the API of the reference's
// ConstructionInfo / build_lattice (llalbm/src/utils/generation/
// ConstructionInfo.hpp:147-503, ConstructionTools.hpp:54-201).  The
// reference keeps std::set's of points and builds the fluid list by
// linear search (quadratic in the lattice size); here every method writes
// a node-class plane directly, so construction is linear and works at
// 8192 x 8192.  Overlaps resolve by the last writer, as in the reference.
// Deviation: a node marked OBSTACLE and later covered by an interval (or
// the other way round) holds only the last class here; the reference
// would keep it in both sets.
#ifndef LLALBM_HIP_CONSTRUCTION_HPP
#define LLALBM_HIP_CONSTRUCTION_HPP

#include <random>
#include <set>
#include <string>
#include <vector>

#include "aliases.hpp"
#include "io.hpp"

namespace llalbm::util::generation
{

enum NonFluidNodeType { BOUNDARY = 0, INLET = 1, OUTLET = 2 };

template <std::size_t dim = 2, std::size_t q = 9>
class ConstructionInfo;

template <>
class ConstructionInfo<2, 9>
{
public:
    ConstructionInfo(std::ostream &out_stream = std::cout) : out(&out_stream) {}

    Index attach_domain_dimensions(const std::array<Index, 2> &dims)
    {
        domain = dims;
        provided = dims[0] > 0 && dims[1] > 0;
        cls.assign(provided ? (std::size_t)(dims[0] * dims[1]) : 0, FLUID);
        return provided ? dims[0] * dims[1] : 0;
    }

    /// ConstructionInfo.hpp:147-231: a straight interval of one node type.
    Index add_nodes_interval(const Point<2> &begin, const Point<2> &end, const NonFluidNodeType &type)
    {
        if (!provided || !in_domain(begin) || !in_domain(end)) return 0;
        std::size_t vary = 0;
        int differing = 0;
        for (std::size_t k = 0; k < 2; ++k)
            if (begin.coords[k] != end.coords[k]) { vary = k; ++differing; }
        if (differing > 1) return 0;
        Point<2> p = begin;
        Index added = 0;
        for (; p.coords[vary] <= end.coords[vary]; ++p.coords[vary], ++added) set(p.coords[0], p.coords[1], to_class(type));
        return added;
    }

    /// ConstructionInfo.hpp:239-281.
    Index add_perimeter_nodes(const NonFluidNodeType &type)
    {
        if (!provided) return 0;
        Index added = 0;
        for (Index j = 0; j < domain[1]; ++j) { set(0, j, to_class(type)); set(domain[0] - 1, j, to_class(type)); added += 2; }
        for (Index i = 0; i < domain[0]; ++i) { set(i, 0, to_class(type)); set(i, domain[1] - 1, to_class(type)); added += 2; }
        return added;
    }

    /// ConstructionInfo.hpp:299-347 (origin = top-left corner, extents
    /// inclusive); as in the reference the (0, 0) point is never an obstacle.
    Index add_obstacle_hyper_rectangle(const Point<2> &origin, const std::array<Index, 2> &ext)
    {
        if (!provided || !in_domain(origin)) return 0;
        Point<2> end{origin.coords[0] + ext[0], origin.coords[1] + ext[1]};
        if (!in_obstacle_domain(end)) return 0;
        Index added = 0;
        for (Index i = origin.coords[0]; i < origin.coords[0] + ext[0]; ++i)
            for (Index j = origin.coords[1]; j < origin.coords[1] + ext[1]; ++j) {
                if (i == 0 && j == 0) continue;
                set(i, j, OBSTACLE);
                ++added;
            }
        return added;
    }

    Index add_obstacle_hyper_square(const Point<2> &origin, const Index &extension)
    {
        return add_obstacle_hyper_rectangle(origin, {extension, extension});
    }

    /// ConstructionInfo.hpp:374-451: integer disc (di^2 + dj^2 <= r^2).
    Index add_obstacle_hyper_sphere(const Point<2> &c, const Index &radius)
    {
        if (!provided || !in_domain(c)) return 0;
        for (std::size_t k = 0; k < 2; ++k) {
            Point<2> hi = c, lo = c;
            hi.coords[k] += radius;
            lo.coords[k] -= radius;
            if (!in_obstacle_domain(hi) || !in_obstacle_domain(lo)) return 0;
        }
        Index added = 0;
        for (Index i = c.coords[0] - radius; i <= c.coords[0] + radius; ++i)
            for (Index j = c.coords[1] - radius; j <= c.coords[1] + radius; ++j) {
                Index di = i - c.coords[0], dj = j - c.coords[1];
                if (di * di + dj * dj <= radius * radius) { set(i, j, OBSTACLE); ++added; }
            }
        return added;
    }

    /// ConstructionInfo.hpp:462-486, same generator and distributions.
    Index add_obstacle_random_spheres(const Index &spheres = 10, const Index &max_radius = 10,
                                      const Index &min_radius = 1, const Index &seed = 42)
    {
        std::mt19937 gen((unsigned)seed);
        std::array<std::uniform_int_distribution<Index>, 2> dists;
        for (std::size_t k = 0; k < 2; ++k) dists[k] = std::uniform_int_distribution<Index>(0, domain[k]);
        std::uniform_int_distribution<Index> dist_radius(min_radius, max_radius);
        Index added = 0;
        for (Index s = 0; s < spheres; ++s) {
            Point<2> c;
            for (std::size_t k = 0; k < 2; ++k) c.coords[k] = dists[k](gen);
            Index r = dist_radius(gen);
            added += add_obstacle_hyper_sphere(c, r);
        }
        return added;
    }

    /// ConstructionInfo.hpp:493-503.
    void read_obstacle_from_file(const std::string &file_path)
    {
        if (!provided) return;
        reader::read_obstacle_classes(file_path, domain, cls);
    }

    std::array<Index, 2> get_domain_dimensions() const { return domain; }
    bool are_dimensions_provided() const { return provided; }
    const std::vector<uint8_t> &node_classes() const { return cls; }

    std::set<BoundaryPoint<2>> get_boundary_nodes() const { return collect<BoundaryPoint<2>>(BOUNDARY_CLS); }
    std::set<BoundaryPoint<2>> get_inlet_nodes() const { return collect<BoundaryPoint<2>>(INLET); }
    std::set<BoundaryPoint<2>> get_outlet_nodes() const { return collect<BoundaryPoint<2>>(OUTLET); }
    std::set<ObstaclePoint<2>> get_obstacle_nodes() const { return collect<ObstaclePoint<2>>(OBSTACLE); }

private:
    static constexpr uint8_t BOUNDARY_CLS = InputNodeType::BOUNDARY;
    static uint8_t to_class(NonFluidNodeType t)
    {
        return t == INLET ? (uint8_t)InputNodeType::INLET : t == OUTLET ? (uint8_t)InputNodeType::OUTLET : BOUNDARY_CLS;
    }
    void set(Index i, Index j, uint8_t k) { cls[(std::size_t)(i + domain[0] * j)] = k; }
    bool in_domain(const Point<2> &p) const
    {
        return p.coords[0] >= 0 && p.coords[0] < domain[0] && p.coords[1] >= 0 && p.coords[1] < domain[1];
    }
    bool in_obstacle_domain(const Point<2> &p) const
    {
        return p.coords[0] > 0 && p.coords[0] < domain[0] - 1 && p.coords[1] > 0 && p.coords[1] < domain[1] - 1;
    }
    template <typename P>
    std::set<P> collect(uint8_t k) const
    {
        std::set<P> s;
        for (Index j = 0; j < domain[1]; ++j)
            for (Index i = 0; i < domain[0]; ++i)
                if (cls[(std::size_t)(i + domain[0] * j)] == k) s.insert(P({i, j}));
        return s;
    }

    std::ostream *out;
    std::array<Index, 2> domain{0, 0};
    bool provided = false;
    std::vector<uint8_t> cls;
};

/// ConstructionTools.hpp:54-201: installs the geometry in the lattice.
template <std::size_t dim, typename Parallelization, template <typename> class LatticeT>
void build_lattice(LatticeT<Parallelization> &lattice, const std::size_t &q, const ConstructionInfo<2, 9> &info)
{
    if (!info.are_dimensions_provided()) throw std::runtime_error("domain dimensions are not provided");
    lattice.reinit_from_classes(q, info.get_domain_dimensions(), info.node_classes());
}

} // namespace llalbm::util::generation

#endif // LLALBM_HIP_CONSTRUCTION_HPP
