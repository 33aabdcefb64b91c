// CPU oracle for the spectral placement step — TEST INFRASTRUCTURE ONLY.
//
// This library restates, on the host, the reference algorithm of each function
// on the placement hot path so the HIP implementation can be checked against it.
// Only tests, the smoke check and the CPU-baseline leg of the benchmark may load
// it; the product path never links or calls it.
//
// Every function names the reference file:line it follows (paths are relative
// to the reference repository root).  Build flags: -ffp-contract=off
// -fno-fast-math, so every double operation rounds exactly once, as in the
// reference.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace sporacle {

using Vec3 = std::array<double, 3>;
using Quat = std::array<double, 4>;  // x, y, z, w
using Tri = std::array<std::uint32_t, 3>;

struct Mesh {
    std::vector<Vec3> vertices;
    std::vector<Tri> triangles;
};

// engine/pack_geometry/src/raster_execution_internal.hpp:9-12 (RasterInterval)
struct Interval {
    double low {}, high {};
    bool valid {};
};
using IVec3 = std::array<Interval, 3>;

// engine/pack_geometry/include/spectrapack/geometry/conservative_fields.hpp:22-31
struct Window {
    Vec3 origin {};
    double pitch {};
    std::array<std::int64_t, 3> first {};
    std::array<std::uint32_t, 3> shape {};
    std::uint64_t cells() const { return std::uint64_t(shape[0]) * shape[1] * shape[2]; }
};

enum class Decision { no, yes, indeterminate };
enum class Sign { negative = -1, zero = 0, positive = 1 };

// ---- interval arithmetic: engine/pack_geometry/src/validation_kernel.cpp:291-324
Interval point_interval(double v);
Interval widened(double low, double high);
Interval add_interval(Interval a, Interval b);
Interval subtract_interval(Interval a, Interval b);
Interval multiply_interval(Interval a, Interval b);
Interval divide_interval(Interval n, Interval d);
std::array<std::array<Interval, 3>, 3> rotation_intervals(const Quat& q);  // :490-508
Interval transformed_interval(const Vec3& p, const Vec3& t, const std::array<std::array<Interval, 3>, 3>& r,
                              int row);                                   // :510-517
bool triangle_box_disjoint(const std::array<IVec3, 3>& tri, const IVec3& box);  // :2054-2079
bool grid_cell_interval(const Window& w, const std::array<std::int64_t, 3>& idx, IVec3& out);  // :2081-2100

// ---- exact predicates: engine/pack_geometry/src/exact_predicates.cpp:236-300,562-745,1056-1110
Sign orient3d(const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d);
Sign orient2d(const std::array<double, 2>& a, const std::array<double, 2>& b, const std::array<double, 2>& c);
enum class Crossing { none, crossing, boundary, degenerate };
Crossing segment_triangle_crossing(const Vec3& p, const Vec3& q, const std::array<Vec3, 3>& tri);
// exact sign of origin + pitch*index + sum(sign_i * v_i): validation_kernel.cpp:2102-2142
int exact_grid_sign(double origin, double pitch, std::int64_t index, const double* values, const int* signs, int n);

// ---- placement of a solid: validation_kernel.cpp:829-943 (translation restricted to 0)
struct Placed {
    const Mesh* mesh {};
    Quat q {};
    bool is_cuboid {}, is_cardinal {}, represented_world_exact {};
    std::array<int, 3> source_axis {}, sign {};
    Vec3 cuboid_min {}, cuboid_max {};  // world values of the cuboid faces
    Vec3 mesh_min {}, mesh_max {};      // AcceptedSolid::bounds_mm of the local mesh
    Vec3 bounds_min {}, bounds_max {};  // conservative (outward) world bounds
    std::vector<Vec3> world;            // represented world vertices (cardinal only)
    std::vector<IVec3> intervals;       // outward vertex intervals
};
bool certify_cuboid(const Mesh& m, Vec3& lo, Vec3& hi);   // validation_kernel.cpp:380-438
bool cardinal_rotation(const Quat& q, std::array<int, 3>& src, std::array<int, 3>& sgn);  // :446-488
bool place(const Mesh& m, const Quat& q, Placed& out);

// ---- fields: engine/pack_geometry/src/conservative_fields.cpp
// Tags: 0 exterior, 1 material, 2 boundary, 3 indeterminate.
bool trim_window(const Placed& p, const Vec3& origin, double pitch, Window& out);  // :506-519, :950-960
void rasterize_boundary(const Placed& p, const Window& w, std::vector<std::uint8_t>& tags);  // vk:3248-3340,3668-3762
Decision classify_material_witness(const Placed& p, const Vec3& wmin, const Vec3& wmax);  // vk:3401-3438
void fill_components(const Placed& p, const Window& w, std::vector<std::uint8_t>& tags,
                     std::uint64_t* components = nullptr, std::uint64_t* uncertain = nullptr);  // cf:257-360
// voxelize_object (cf:919-1003): {0,1} kernel cells.
bool voxelize_object(const Mesh& m, const Vec3& origin, double pitch, const Quat& q, Window& w,
                     std::vector<std::uint8_t>& cells, std::vector<std::uint8_t>* raw_tags = nullptr);
// voxelize_container, analytic box (cf:1146-1198): 1 = blocked, 0 = certified free.
bool voxelize_box_container(const Window& w, const Vec3& dims, double clearance, std::vector<std::uint8_t>& cells);
// voxelize_container, mesh (cf:1200-1275) at zero clearance, or with a Euclidean stencil.
bool voxelize_mesh_container(const Mesh& m, const Window& w, double clearance, std::vector<std::uint8_t>& cells);
// euclidean_offset_reaches (vk:3492-3522)
Decision euclidean_offset_reaches(const std::array<std::int64_t, 3>& d, double pitch, double clearance);

// ---- orientation catalogs: engine/pack_solver/src/orientation_cube.hpp:10-80, orientations.cpp:44-172
std::array<Quat, 24> cube_seed_array();
Quat canonical(Quat q);
bool valid_quaternion(const Quat& q);
std::vector<Quat> custom_catalog(const std::vector<Quat>& raw);
std::array<Quat, 96> cube45_orientations();

// ---- correlation: engine/pack_compute/src/correlation.cpp:238-655
struct CorrelationSpec {
    std::array<std::uint32_t, 3> environment_shape {}, kernel_shape {};
    std::array<std::int64_t, 3> environment_first {}, kernel_first {};
};
struct CorrelationLimits {
    std::uint64_t max_working_bytes { 512ULL << 20 };
    std::uint64_t reserved_bytes {};
    std::uint64_t max_padded_cells { 16777216 };
    std::uint64_t max_direct_terms { 200000000 };
};
struct CorrelationOut {
    std::string code;  // empty on success
    std::array<std::int64_t, 3> translation_first {};
    std::array<std::uint32_t, 3> shape {};
    std::vector<double> values;
    double max_integer_residual {}, mass_residual {}, max_probe_error {};
    std::uint64_t padded_cells {}, working_bytes_peak {}, direct_terms {};
};
// Numeric fault seams of correlation_test_hook.hpp:6-10.
enum class NumericFault { none, bad_normalization, nan, integral_corruption };
CorrelationOut correlate_binary(const CorrelationSpec& s, const std::uint8_t* env, std::uint64_t env_n,
                                const std::uint8_t* ker, std::uint64_t ker_n, const CorrelationLimits& l,
                                NumericFault fault = NumericFault::none);
CorrelationOut correlate_proximity(const CorrelationSpec& s, const double* env, std::uint64_t env_n,
                                   const std::uint8_t* ker, std::uint64_t ker_n, const CorrelationLimits& l,
                                   NumericFault fault = NumericFault::none);
struct Estimate {
    std::string code;
    std::array<std::uint32_t, 3> padded_shape {};
    std::uint64_t padded_cells {}, working_bytes {};
};
Estimate estimate_correlation(const CorrelationSpec& s, std::uint32_t fft_threads);  // correlation.cpp:596-641

// tests/support/src/integer_correlation.cpp:76-130
struct DirectOut {
    std::string code;
    std::array<int, 6> range {};  // min_x, max_x, min_y, max_y, min_z, max_z
    std::array<int, 3> extent {};
    std::vector<int> values;
};
DirectOut direct_linear_cross_correlation(const std::vector<int>& blocked, std::array<int, 3> b_ext,
                                          const std::vector<int>& kernel, std::array<int, 3> k_ext,
                                          std::array<int, 3> origin);

// ---- FP64 complex FFT along one axis (the role pocketfft::c2c plays in fft_execution.cpp:74)
void fft_axis(std::complex<double>* data, const std::array<std::uint32_t, 3>& shape, int axis, bool forward,
              double factor);

// ---- plain placement step: engine/pack_solver/src/gap_fill.cpp:363-509
struct KernelField {
    std::array<std::int64_t, 3> first {};
    std::array<std::uint32_t, 3> shape {};
    std::vector<std::uint8_t> cells;
};
struct StepChoice {
    int orientation { -1 };
    std::array<std::int64_t, 3> translation {};
    std::array<std::int64_t, 3> key {};
};
// Free mask of one orientation over legal translations (X-fastest), from
// an overlap field; returns false when the kernel does not fit.
bool free_mask(const std::array<std::int64_t, 3>& env_first, const std::array<std::uint32_t, 3>& env_shape,
               const KernelField& k, const CorrelationOut& c, std::vector<std::uint8_t>& mask,
               std::array<std::int64_t, 3>& low, std::array<std::uint32_t, 3>& legal);
// One step: correlate every orientation (direct=false uses the FFT restatement,
// direct=true the integer oracle), then the lowest (z,y,x) rule.
std::string place_step(const std::vector<std::uint8_t>& env, const std::array<std::int64_t, 3>& env_first,
                       const std::array<std::uint32_t, 3>& env_shape, const std::vector<KernelField>& kernels,
                       bool direct, StepChoice& choice, std::vector<std::vector<std::int32_t>>* legal_fields = nullptr,
                       const std::vector<std::pair<int, std::array<std::int64_t, 3>>>* excluded = nullptr);
// gap_fill.cpp:98-121 for a kernel at translation t.
void stamp(std::vector<std::uint8_t>& env, const std::array<std::int64_t, 3>& env_first,
           const std::array<std::uint32_t, 3>& env_shape, const KernelField& k, const std::array<std::int64_t, 3>& t);

// ---- proximity field: engine/pack_solver/src/field_proximity.hpp:32-132
void build_proximity(const std::vector<std::uint8_t>& occupied, const std::array<std::uint32_t, 3>& shape,
                     std::vector<double>& proximity, std::vector<std::uint64_t>* squared = nullptr);

}  // namespace sporacle
