// Oracle (test infrastructure): plain C entry points for the tests, the smoke
// check and the CPU-baseline leg of the benchmark (loaded through ctypes).
#include <cstring>
#include <thread>

#include <cmath>

#include "sp_oracle.hpp"

using namespace sporacle;

namespace {
void put_code(const std::string& code, char* out, int len)
{
    if (!out || len <= 0) return;
    std::strncpy(out, code.c_str(), std::size_t(len - 1));
    out[len - 1] = 0;
}
Mesh make_mesh(const double* v, std::uint32_t nv, const std::uint32_t* t, std::uint32_t nt)
{
    Mesh m;
    m.vertices.resize(nv);
    for (std::uint32_t i = 0; i != nv; ++i) m.vertices[i] = { v[3 * i], v[3 * i + 1], v[3 * i + 2] };
    m.triangles.resize(nt);
    for (std::uint32_t i = 0; i != nt; ++i) m.triangles[i] = { t[3 * i], t[3 * i + 1], t[3 * i + 2] };
    return m;
}
}  // namespace

extern "C" {

int orc_direct_correlation(const int* b, const int* be, const int* k, const int* ke, const int* origin, int* range,
                           int* extent, int* values, std::uint64_t capacity)
{
    const std::array<int, 3> bex { be[0], be[1], be[2] }, kex { ke[0], ke[1], ke[2] }, org { origin[0], origin[1], origin[2] };
    std::vector<int> bb, kk;
    if (bex[0] > 0 && bex[1] > 0 && bex[2] > 0) bb.assign(b, b + std::size_t(bex[0]) * bex[1] * bex[2]);
    if (kex[0] > 0 && kex[1] > 0 && kex[2] > 0) kk.assign(k, k + std::size_t(kex[0]) * kex[1] * kex[2]);
    const DirectOut o = direct_linear_cross_correlation(bb, bex, kk, kex, org);
    if (o.code == "invalid_argument") return 1;
    if (o.code == "overflow_error") return 2;
    for (int i = 0; i != 6; ++i) range[i] = o.range[i];
    for (int i = 0; i != 3; ++i) extent[i] = o.extent[i];
    if (o.values.size() > capacity) return 3;
    std::memcpy(values, o.values.data(), o.values.size() * sizeof(int));
    return 0;
}

static int finish(const CorrelationOut& o, std::int64_t* out_first, std::uint32_t* out_shape, double* values,
                  std::uint64_t capacity, double* report, std::uint64_t* stats, char* code, int code_len)
{
    stats[0] = o.padded_cells;
    stats[1] = o.working_bytes_peak;
    stats[2] = o.direct_terms;
    put_code(o.code, code, code_len);
    if (!o.code.empty()) return 1;
    for (int a = 0; a != 3; ++a) {
        out_first[a] = o.translation_first[a];
        out_shape[a] = o.shape[a];
    }
    report[0] = o.max_integer_residual;
    report[1] = o.mass_residual;
    report[2] = o.max_probe_error;
    if (o.values.size() > capacity) return 2;
    std::memcpy(values, o.values.data(), o.values.size() * sizeof(double));
    return 0;
}

static CorrelationSpec spec_of(const std::uint32_t* es, const std::uint32_t* ks, const std::int64_t* ef,
                               const std::int64_t* kf)
{
    return { { es[0], es[1], es[2] }, { ks[0], ks[1], ks[2] }, { ef[0], ef[1], ef[2] }, { kf[0], kf[1], kf[2] } };
}

int orc_correlate_binary(const std::uint32_t* es, const std::uint32_t* ks, const std::int64_t* ef, const std::int64_t* kf,
                         const std::uint8_t* env, std::uint64_t env_n, const std::uint8_t* ker, std::uint64_t ker_n,
                         const std::uint64_t* limits, int fault, std::int64_t* out_first, std::uint32_t* out_shape,
                         double* values, std::uint64_t capacity, double* report, std::uint64_t* stats, char* code,
                         int code_len)
{
    const CorrelationLimits l { limits[0], limits[1], limits[2], limits[3] };
    const auto o = correlate_binary(spec_of(es, ks, ef, kf), env, env_n, ker, ker_n, l, NumericFault(fault));
    return finish(o, out_first, out_shape, values, capacity, report, stats, code, code_len);
}

int orc_correlate_proximity(const std::uint32_t* es, const std::uint32_t* ks, const std::int64_t* ef,
                            const std::int64_t* kf, const double* env, std::uint64_t env_n, const std::uint8_t* ker,
                            std::uint64_t ker_n, const std::uint64_t* limits, int fault, std::int64_t* out_first,
                            std::uint32_t* out_shape, double* values, std::uint64_t capacity, double* report,
                            std::uint64_t* stats, char* code, int code_len)
{
    const CorrelationLimits l { limits[0], limits[1], limits[2], limits[3] };
    const auto o = correlate_proximity(spec_of(es, ks, ef, kf), env, env_n, ker, ker_n, l, NumericFault(fault));
    return finish(o, out_first, out_shape, values, capacity, report, stats, code, code_len);
}

int orc_estimate(const std::uint32_t* es, const std::uint32_t* ks, const std::int64_t* ef, const std::int64_t* kf,
                 std::uint32_t threads, std::uint32_t* padded_shape, std::uint64_t* cells_bytes, char* code, int code_len)
{
    const auto e = estimate_correlation(spec_of(es, ks, ef, kf), threads);
    put_code(e.code, code, code_len);
    if (!e.code.empty()) return 1;
    for (int a = 0; a != 3; ++a) padded_shape[a] = e.padded_shape[a];
    cells_bytes[0] = e.padded_cells;
    cells_bytes[1] = e.working_bytes;
    return 0;
}

void orc_fft_1d(double* interleaved, std::uint32_t n, int forward)
{
    fft_axis(reinterpret_cast<std::complex<double>*>(interleaved), { n, 1, 1 }, 0, forward != 0, 1.0);
}

void orc_cube_catalog(double* out)
{
    const auto c = cube_seed_array();
    for (int i = 0; i != 24; ++i)
        for (int j = 0; j != 4; ++j) out[4 * i + j] = c[i][j];
}

void orc_cube45_catalog(double* out)
{
    const auto c = cube45_orientations();
    for (int i = 0; i != 96; ++i)
        for (int j = 0; j != 4; ++j) out[4 * i + j] = c[i][j];
}

int orc_custom_catalog(const double* in, std::uint32_t n, double* out, std::uint32_t* count)
{
    std::vector<Quat> raw(n);
    for (std::uint32_t i = 0; i != n; ++i) raw[i] = { in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3] };
    const auto c = custom_catalog(raw);
    if (c.empty()) return 1;
    *count = std::uint32_t(c.size());
    for (std::size_t i = 0; i != c.size(); ++i)
        for (int j = 0; j != 4; ++j) out[4 * i + j] = c[i][j];
    return 0;
}

// Returns 0 ok, 1 failure, 2 capacity too small (first/shape filled).
int orc_voxelize_object(const double* v, std::uint32_t nv, const std::uint32_t* t, std::uint32_t nt, const double* origin,
                        double pitch, const double* q, std::int64_t* first, std::uint32_t* shape, std::uint8_t* cells,
                        std::uint64_t capacity, std::uint8_t* raw_tags)
{
    const Mesh m = make_mesh(v, nv, t, nt);
    Window w;
    std::vector<std::uint8_t> c, raw;
    if (!voxelize_object(m, { origin[0], origin[1], origin[2] }, pitch, { q[0], q[1], q[2], q[3] }, w, c,
                         raw_tags ? &raw : nullptr))
        return 1;
    for (int a = 0; a != 3; ++a) {
        first[a] = w.first[a];
        shape[a] = w.shape[a];
    }
    if (c.size() > capacity) return 2;
    std::memcpy(cells, c.data(), c.size());
    if (raw_tags) std::memcpy(raw_tags, raw.data(), raw.size());
    return 0;
}

int orc_object_bounds(const double* v, std::uint32_t nv, const std::uint32_t* t, std::uint32_t nt, const double* q,
                      double* bmin, double* bmax, int* flags)
{
    const Mesh m = make_mesh(v, nv, t, nt);
    Placed p;
    if (!place(m, { q[0], q[1], q[2], q[3] }, p)) return 1;
    for (int a = 0; a != 3; ++a) {
        bmin[a] = p.bounds_min[a];
        bmax[a] = p.bounds_max[a];
    }
    flags[0] = p.is_cuboid;
    flags[1] = p.is_cardinal;
    flags[2] = p.represented_world_exact;
    return 0;
}

int orc_box_container(const std::int64_t* first, const std::uint32_t* shape, const double* origin, double pitch,
                      const double* dims, double clearance, std::uint8_t* cells)
{
    Window w;
    w.origin = { origin[0], origin[1], origin[2] };
    w.pitch = pitch;
    for (int a = 0; a != 3; ++a) {
        w.first[a] = first[a];
        w.shape[a] = shape[a];
    }
    std::vector<std::uint8_t> c;
    if (!voxelize_box_container(w, { dims[0], dims[1], dims[2] }, clearance, c)) return 1;
    std::memcpy(cells, c.data(), c.size());
    return 0;
}

int orc_mesh_container(const double* v, std::uint32_t nv, const std::uint32_t* t, std::uint32_t nt,
                       const std::int64_t* first, const std::uint32_t* shape, const double* origin, double pitch,
                       double clearance, std::uint8_t* cells)
{
    const Mesh m = make_mesh(v, nv, t, nt);
    Window w;
    w.origin = { origin[0], origin[1], origin[2] };
    w.pitch = pitch;
    for (int a = 0; a != 3; ++a) {
        w.first[a] = first[a];
        w.shape[a] = shape[a];
    }
    std::vector<std::uint8_t> c;
    if (!voxelize_mesh_container(m, w, clearance, c)) return 1;
    std::memcpy(cells, c.data(), c.size());
    return 0;
}

static std::vector<KernelField> kernels_of(std::uint32_t n, const std::uint8_t* const* cells, const std::uint32_t* ks,
                                           const std::int64_t* kf)
{
    std::vector<KernelField> v(n);
    for (std::uint32_t i = 0; i != n; ++i) {
        for (int a = 0; a != 3; ++a) {
            v[i].shape[a] = ks[3 * i + a];
            v[i].first[a] = kf[3 * i + a];
        }
        v[i].cells.assign(cells[i], cells[i] + std::size_t(v[i].shape[0]) * v[i].shape[1] * v[i].shape[2]);
    }
    return v;
}

// One plain placement step.  fields (optional) receives each orientation's
// legal overlap counts, concatenated in orientation order.
int orc_place_step(const std::uint8_t* env, const std::int64_t* ef, const std::uint32_t* es, std::uint32_t n,
                   const std::uint8_t* const* kcells, const std::uint32_t* ks, const std::int64_t* kf, int direct,
                   int* orientation, std::int64_t* translation, std::int32_t* fields, char* code, int code_len,
                   const std::int64_t* excluded, std::uint32_t n_excluded)
{
    const std::vector<std::uint8_t> e(env, env + std::size_t(es[0]) * es[1] * es[2]);
    const auto kernels = kernels_of(n, kcells, ks, kf);
    StepChoice c;
    std::vector<std::vector<std::int32_t>> f;
    std::vector<std::pair<int, std::array<std::int64_t, 3>>> ex;
    for (std::uint32_t i = 0; i != n_excluded; ++i)
        ex.push_back({ int(excluded[4 * i]), { excluded[4 * i + 1], excluded[4 * i + 2], excluded[4 * i + 3] } });
    const auto err = place_step(e, { ef[0], ef[1], ef[2] }, { es[0], es[1], es[2] }, kernels, direct != 0, c,
                                fields ? &f : nullptr, ex.empty() ? nullptr : &ex);
    put_code(err, code, code_len);
    if (!err.empty()) return 1;
    *orientation = c.orientation;
    for (int a = 0; a != 3; ++a) translation[a] = c.translation[a];
    if (fields)
        for (const auto& x : f) {
            std::memcpy(fields, x.data(), x.size() * sizeof(std::int32_t));
            fields += x.size();
        }
    return 0;
}

// A bounded CPU-baseline sample: the FFT restatement of one step's
// correlations for orientations [begin, end), over `threads` host threads
// (orientations in parallel).  Returns the number of correlations done.
int orc_correlate_orientations(const std::uint8_t* env, const std::int64_t* ef, const std::uint32_t* es, std::uint32_t n,
                               const std::uint8_t* const* kcells, const std::uint32_t* ks, const std::int64_t* kf,
                               std::uint32_t begin, std::uint32_t end, std::uint32_t threads, std::int64_t* checksum)
{
    const auto kernels = kernels_of(n, kcells, ks, kf);
    const std::size_t env_n = std::size_t(es[0]) * es[1] * es[2];
    CorrelationLimits lim;
    lim.max_working_bytes = ~0ULL >> 1;
    lim.max_padded_cells = ~0ULL >> 1;
    lim.max_direct_terms = ~0ULL;
    std::vector<std::int64_t> sums(end - begin, 0);
    std::vector<int> ok(end - begin, 0);
    auto work = [&](std::uint32_t tid) {
        for (std::uint32_t o = begin + tid; o < end; o += threads) {
            const auto& k = kernels[o];
            const CorrelationSpec s { { es[0], es[1], es[2] }, k.shape, { ef[0], ef[1], ef[2] }, k.first };
            const auto c = correlate_binary(s, env, env_n, k.cells.data(), k.cells.size(), lim);
            if (!c.code.empty()) continue;
            std::int64_t acc = 0;
            for (double v : c.values) acc += std::int64_t(v);
            sums[o - begin] = acc;
            ok[o - begin] = 1;
        }
    };
    std::vector<std::thread> pool;
    for (std::uint32_t i = 1; i < threads; ++i) pool.emplace_back(work, i);
    work(0);
    for (auto& th : pool) th.join();
    int done = 0;
    *checksum = 0;
    for (std::size_t i = 0; i != sums.size(); ++i) {
        done += ok[i];
        *checksum += sums[i];
    }
    return done;
}

void orc_stamp(std::uint8_t* env, const std::int64_t* ef, const std::uint32_t* es, const std::uint8_t* kcells,
               const std::uint32_t* ks, const std::int64_t* kf, const std::int64_t* t)
{
    std::vector<std::uint8_t> e(env, env + std::size_t(es[0]) * es[1] * es[2]);
    KernelField k;
    for (int a = 0; a != 3; ++a) {
        k.shape[a] = ks[a];
        k.first[a] = kf[a];
    }
    k.cells.assign(kcells, kcells + std::size_t(ks[0]) * ks[1] * ks[2]);
    stamp(e, { ef[0], ef[1], ef[2] }, { es[0], es[1], es[2] }, k, { t[0], t[1], t[2] });
    std::memcpy(env, e.data(), e.size());
}

// The pair-clearance stencil (conservative_fields.cpp:189-204 halo,
// 398-434 make_stencil): offsets within the halo that euclidean_offset_reaches
// does not rule out.  out holds up to capacity (x, y, z) triples; returns the
// count (or -1 for an invalid clearance).
int orc_clearance_stencil(double pitch, double clearance, std::int64_t* out, std::uint32_t capacity)
{
    if (!std::isfinite(clearance) || clearance < 0.0) return -1;
    std::int64_t halo = 0;
    if (clearance != 0.0) halo = static_cast<std::int64_t>(std::ceil(static_cast<long double>(clearance) / pitch) + 1.0L);
    int count = 0;
    for (std::int64_t z = -halo; z <= halo; ++z)
        for (std::int64_t y = -halo; y <= halo; ++y)
            for (std::int64_t x = -halo; x <= halo; ++x)
                if (sporacle::euclidean_offset_reaches({ x, y, z }, pitch, clearance) != sporacle::Decision::no) {
                    if (std::uint32_t(count) < capacity) {
                        out[3 * count] = x;
                        out[3 * count + 1] = y;
                        out[3 * count + 2] = z;
                    }
                    ++count;
                }
    return count;
}

void orc_proximity(const std::uint8_t* occ, const std::uint32_t* shape, double* values, std::uint64_t* squared)
{
    const std::vector<std::uint8_t> o(occ, occ + std::size_t(shape[0]) * shape[1] * shape[2]);
    std::vector<double> v;
    std::vector<std::uint64_t> sq;
    build_proximity(o, { shape[0], shape[1], shape[2] }, v, squared ? &sq : nullptr);
    std::memcpy(values, v.data(), v.size() * sizeof(double));
    if (squared) std::memcpy(squared, sq.data(), sq.size() * sizeof(std::uint64_t));
}

int orc_segment_triangle_crossing(const double* p, const double* q, const double* tri)
{
    return int(segment_triangle_crossing({ p[0], p[1], p[2] }, { q[0], q[1], q[2] },
                                         { Vec3 { tri[0], tri[1], tri[2] }, Vec3 { tri[3], tri[4], tri[5] },
                                           Vec3 { tri[6], tri[7], tri[8] } }));
}

int orc_orient3d(const double* a, const double* b, const double* c, const double* d)
{
    return int(orient3d({ a[0], a[1], a[2] }, { b[0], b[1], b[2] }, { c[0], c[1], c[2] }, { d[0], d[1], d[2] }));
}

}  // extern "C"
