// Data model of the LLALBM-compatible host layer: points, boundary types,
// node classes and a small ColMajor host tensor.  Names, enumerator values
// and comparison semantics follow llalbm/src/utils/aliases.hpp:28-357 so
// that user code written against the reference compiles against this layer.
#ifndef LLALBM_HIP_ALIASES_HPP
#define LLALBM_HIP_ALIASES_HPP

#include <algorithm>
#include <array>
#include <bitset>
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <vector>

namespace llalbm::util
{

using Index = std::int64_t;

/// Boundary types (aliases.hpp:28-38).
enum btypes {
    NONE = 0,
    TOP_WALL,
    RIGHT_WALL,
    BOTTOM_WALL,
    LEFT_WALL,
    RIGHT_TOP_C,
    RIGHT_BOTTOM_C,
    LEFT_BOTTOM_C,
    LEFT_TOP_C
};

/// Node types of the lattice file format (aliases.hpp:41-48); the same
/// values are the node classes of the C ABI (lbm_node_class).
enum InputNodeType {
    FLUID = 0,
    SOLID,
    BOUNDARY,
    INLET,
    OUTLET,
    OBSTACLE
};

/// A point of the lattice, coords[0] = row, coords[1] = column.
template <std::size_t d>
struct Point {
    std::array<Index, d> coords{};

    Point() = default;
    Point(const std::array<Index, d> &c) : coords(c) {}
    Point(std::initializer_list<Index> l)
    {
        assert(l.size() == d && "The number of coordinates must match the number of dimensions");
        std::size_t k = 0;
        for (Index v : l) coords[k++] = v;
    }
    bool operator==(const Point &o) const { return coords == o.coords; }
    bool operator!=(const Point &o) const { return !(*this == o); }
    bool operator<(const Point &o) const { return coords < o.coords; }  // lexicographic, row first
};

/// A point on the domain boundary with its type.
template <std::size_t d>
struct BoundaryPoint {
    std::array<Index, d> coords{};
    btypes type = NONE;

    BoundaryPoint() = default;
    BoundaryPoint(const std::array<Index, d> &c) : coords(c), type(NONE) {}
    BoundaryPoint(const Point<d> &p) : coords(p.coords), type(NONE) {}
    BoundaryPoint(std::initializer_list<Index> l)
    {
        assert(l.size() == d && "The number of coordinates must match the number of dimensions");
        std::size_t k = 0;
        for (Index v : l) coords[k++] = v;
    }
    bool operator==(const BoundaryPoint &o) const { return coords == o.coords && type == o.type; }
    bool operator<(const BoundaryPoint &o) const { return coords < o.coords; }
};

template <std::size_t d>
struct ObstaclePoint;

/// Obstacle node with its bounce-back directions: bit k set when the
/// neighbour in D2Q9 direction k is fluid, inlet or outlet.
template <>
struct ObstaclePoint<2> {
    std::array<Index, 2> coords{};
    double b = 1.0;
    std::bitset<9> directions;

    ObstaclePoint() = default;
    ObstaclePoint(const std::array<Index, 2> &c) : coords(c) {}
    ObstaclePoint(const Point<2> &p) : coords(p.coords) {}
    ObstaclePoint(std::initializer_list<Index> l)
    {
        assert(l.size() == 2 && "The number of coordinates must match the number of dimensions");
        std::size_t k = 0;
        for (Index v : l) coords[k++] = v;
    }
    bool operator==(const ObstaclePoint &o) const { return coords == o.coords; }
    bool operator<(const ObstaclePoint &o) const { return coords < o.coords; }
    double get_b() const { return b; }
    void set_b(double nb) { b = nb; }
};

/// Host tensor in ColMajor layout: element (i, j, k) of an (n0, n1, n2)
/// tensor at i + n0*(j + n1*k), the layout of the reference's
/// Eigen::Tensor (and of the C ABI's host arrays).
template <typename T, std::size_t N>
class HostTensor
{
public:
    HostTensor() { dims_.fill(0); }
    explicit HostTensor(const std::array<Index, N> &dims) : dims_(dims), data_(count(dims)) {}
    template <typename... I>
    explicit HostTensor(Index first, I... rest) : HostTensor(std::array<Index, N>{first, (Index)rest...})
    {
    }

    const std::array<Index, N> &dimensions() const { return dims_; }
    Index dimension(std::size_t k) const { return dims_[k]; }
    Index size() const { return (Index)data_.size(); }
    T *data() { return data_.data(); }
    const T *data() const { return data_.data(); }
    void setConstant(const T &v) { std::fill(data_.begin(), data_.end(), v); }

    template <typename... I>
    T &operator()(I... idx) { return data_[offset({(Index)idx...})]; }
    template <typename... I>
    const T &operator()(I... idx) const { return data_[offset({(Index)idx...})]; }
    T &operator()(const std::array<Index, N> &idx) { return data_[offset(idx)]; }
    const T &operator()(const std::array<Index, N> &idx) const { return data_[offset(idx)]; }

private:
    static std::size_t count(const std::array<Index, N> &d)
    {
        std::size_t n = 1;
        for (Index v : d) n *= (std::size_t)v;
        return n;
    }
    std::size_t offset(const std::array<Index, N> &idx) const
    {
        std::size_t off = 0;
        for (std::size_t k = N; k-- > 0;) off = off * (std::size_t)dims_[k] + (std::size_t)idx[k];
        return off;
    }
    std::array<Index, N> dims_;
    std::vector<T> data_;
};

} // namespace llalbm::util

namespace llalbm
{
template <typename T, std::size_t N>
using Tensor = util::HostTensor<T, N>;
} // namespace llalbm

#endif // LLALBM_HIP_ALIASES_HPP
