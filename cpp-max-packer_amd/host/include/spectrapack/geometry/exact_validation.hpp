// Exact host validation of final placements: the role of
// engine/pack_geometry/src/validation.cpp:517-634 (validate: pair solids, then
// containment) for the meshes this build packs.
//
// A placed copy is the exact image of the part mesh under the rotation of its
// quaternion (validation_kernel.cpp:490-517: the matrix H(q) / |q|^2, which is
// a rotation for any nonzero q) and its translation, both taken as the exact
// values of their doubles.  Every predicate (orient3d, orient2d, coordinate
// comparisons) is decided on those exact values: an interval evaluation
// decides it when its sign is certain, and an exact big-number evaluation
// otherwise, so no answer depends on rounding.
//
// Verdicts follow the reference's tri-state (valid / invalid /
// indeterminate) and codes:
//   VALID                               every pair and every containment certified
//   VALIDATION_PAIR                     two copies' interiors overlap (an edge of one
//                                       crosses a face of the other strictly inside it,
//                                       or one copy lies inside the other)
//   VALIDATION_PAIR_UNRESOLVED          surfaces touch without crossing, or a pair
//                                       clearance above zero that the bounding boxes
//                                       do not certify
//   VALIDATION_CONTAINER                a copy reaches outside the container (box: past
//                                       the walls shrunk by the wall clearance; mesh: a
//                                       crossing of the container surface, or outside)
//   VALIDATION_CONTAINER_UNRESOLVED     the copy touches the mesh container's surface,
//                                       or a mesh container with a wall clearance
// A pair is certified disjoint when the exact axis-aligned bounds of the two
// copies are separated by at least the pair clearance on some axis
// (bounds_pair_certificate's role), or, with no clearance, when no triangle of
// one meets (closed) a triangle of the other and neither copy holds a vertex
// of the other.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <span>
#include <string_view>
#include <variant>
#include <vector>

namespace spectrapack::geometry::exact {

struct Mesh {
    std::vector<std::array<double, 3>> vertices;
    std::vector<std::array<std::uint32_t, 3>> triangles;  // closed, consistently oriented
};

struct Pose {
    std::array<double, 3> translation_mm {};
    std::array<double, 4> rotation_xyzw { 0, 0, 0, 1 };
};

// Box container [0, dims] per axis (BoxDimensions); a closed mesh container
// whose interior is the free region (world frame).
struct BoxContainer {
    std::array<double, 3> dims_mm {};
};
using Container = std::variant<BoxContainer, const Mesh*>;

enum class Validity { valid, invalid, indeterminate };

constexpr std::size_t kNoCopy = std::numeric_limits<std::size_t>::max();

struct ValidationReport {
    Validity validity { Validity::indeterminate };
    std::string_view code { "VALIDATION_NOT_RUN" };
    std::size_t first { kNoCopy }, second { kNoCopy };  // affected copies (indices into the poses)
    std::uint64_t pair_tests {};        // copy pairs whose bounds overlap
    std::uint64_t triangle_tests {};    // triangle pairs classified
    std::uint64_t exact_evaluations {}; // predicates the interval filter left to exact arithmetic
};

// validate (validation.cpp:517-634): every pair of copies, then every copy
// against the container.
ValidationReport validate_placements(const Mesh& part, std::span<const Pose> poses, const Container& container,
                                     double pair_clearance_mm = 0.0, double wall_clearance_mm = 0.0);

// One more copy against copies already validated together and the container:
// what a placement loop's accept verdict needs (validate_incremental's role,
// validation.cpp:530-560, with every earlier pose proven).
class Validator {
public:
    Validator(const Mesh& part, Container container, double pair_clearance_mm = 0.0, double wall_clearance_mm = 0.0);
    ~Validator();
    Validator(const Validator&) = delete;
    Validator& operator=(const Validator&) = delete;
    // Classifies `pose` against the container and every accepted copy; the
    // copy is kept (accepted) only when the verdict is valid.
    ValidationReport try_add(const Pose& pose);
    std::size_t accepted() const;

private:
    struct State;
    State* state_;
    friend ValidationReport validate_placements(const Mesh&, std::span<const Pose>, const Container&, double, double);
};

}  // namespace spectrapack::geometry::exact
