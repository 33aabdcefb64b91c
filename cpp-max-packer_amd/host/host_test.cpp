// Test driver of the C++ host mirror (run by tests/test_gpu_host.py).
//   spectrapack_host_test correlate          — golden vectors, failure variants and admission
//   spectrapack_host_test estimate           — the admission estimate alone (no device work)
//   spectrapack_host_test place <input.txt>  — the placement loop on a scene
//     read from whitespace-separated numbers; prints one line per proposal
//     "P orientation tx ty tz verdict" and a final "S steps proposals accepted outcome".
//   spectrapack_host_test place-exact <input.txt> — the placement loop with the
//     exact host validation as its accept verdict
//   spectrapack_host_test rank <input.txt>   — the ranked loop on a scene (the
//     same numbers, then origin_z cmin_z cmax_z, one max z per orientation,
//     capacity and reject_every); prints "R orientation tx ty tz score_bits"
//     per placement and a final "S placed validations outcome mass".
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <utility>
#include <stop_token>
#include <string>
#include <vector>

#include "spectrapack/compute/correlation.hpp"
#include "spectrapack/solver/placement.hpp"

namespace sc = spectrapack::compute;
namespace sg = spectrapack::solver::gpu;

static int failures = 0;
#define CHECK(cond)                                                                     \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);                  \
            ++failures;                                                                 \
        }                                                                               \
    } while (0)

static void admission_checks();

static int correlate_checks()
{
    // engine/pack_compute/tests/correlation_tests.cpp:292-308
    const sc::CorrelationSpec spec { { 3, 2, 1 }, { 2, 2, 1 }, { 4, -3, 0 }, { -1, 1, 0 } };
    const std::vector<std::uint8_t> env { 1, 0, 1, 0, 1, 1 };
    const std::vector<std::uint8_t> ker { 1, 0, 1, 1 };
    const auto out = sc::correlate_binary_cpu(spec, env, ker);
    CHECK(std::holds_alternative<sc::CorrelationResult>(out));
    if (const auto* r = std::get_if<sc::CorrelationResult>(&out)) {
        CHECK((r->translation_first == sc::Index3 { 4, -5, 0 }));
        CHECK((r->shape == sc::Shape3 { 4, 3, 1 }));
        const std::vector<double> expected { 1, 1, 1, 1, 0, 2, 2, 2, 0, 0, 1, 1 };
        CHECK(r->values == expected);
        CHECK(r->stats.padded_cells == 12);
    }
    // failure variants carry the reference's code strings
    const std::vector<std::uint8_t> bad { 1, 0, 2, 1 };
    const auto f1 = sc::correlate_binary_cpu(spec, env, bad);
    CHECK(std::holds_alternative<sc::CorrelationFailure>(f1) &&
          std::get<sc::CorrelationFailure>(f1).code == "CORRELATION_INVALID_KERNEL");
    sc::CorrelationLimits tight;
    tight.max_padded_cells = 11;
    const auto f2 = sc::correlate_binary_cpu(spec, env, ker, tight);
    CHECK(std::holds_alternative<sc::CorrelationFailure>(f2) &&
          std::get<sc::CorrelationFailure>(f2).code == "CORRELATION_PADDED_CELL_LIMIT");
    std::stop_source stop;
    stop.request_stop();
    spectrapack::runtime::OperationControl stopped;
    stopped.stop = stop.get_token();
    const auto f3 = sc::correlate_binary_cpu(spec, env, ker, {}, stopped);
    CHECK(std::holds_alternative<sc::CorrelationFailure>(f3) &&
          std::get<sc::CorrelationFailure>(f3).code == "OPERATION_CANCELLED");
    spectrapack::runtime::OperationControl late;
    late.deadline = spectrapack::runtime::Clock::now() - std::chrono::seconds(1);
    const auto f4 = sc::correlate_binary_cpu(spec, env, ker, {}, late);
    CHECK(std::holds_alternative<sc::CorrelationFailure>(f4) &&
          std::get<sc::CorrelationFailure>(f4).code == "DEADLINE_EXCEEDED");
    const auto e = sc::estimate_correlation_cpu(spec, 9);
    CHECK(std::holds_alternative<sc::CorrelationFailure>(e) &&
          std::get<sc::CorrelationFailure>(e).code == "CPU_THREAD_COUNT_UNSUPPORTED");
    const auto e2 = sc::estimate_correlation_cpu(spec, 1);
    CHECK(std::holds_alternative<sc::CorrelationEstimate>(e2) &&
          std::get<sc::CorrelationEstimate>(e2).padded_cells == 12);
    // proximity keeps fractional values
    const std::vector<double> penv { 0.5, 0.25, 1.0, 0.0, 0.75, 0.125 };
    const auto p = sc::correlate_proximity_cpu(spec, penv, ker);
    CHECK(std::holds_alternative<sc::CorrelationResult>(p));
    if (const auto* r = std::get_if<sc::CorrelationResult>(&p)) CHECK(std::abs(r->values[5] - 1.25) < 1e-12);
    admission_checks();
    std::printf("correlate checks: %d failures\n", failures);
    return failures == 0 ? 0 : 1;
}

// engine/pack_compute/tests/correlation_tests.cpp:200-214: the admission
// bytes the reference expects for a padded cell count and longest axis.
static std::uint64_t metadata_bytes()
{
    return 2 * sizeof(sc::CorrelationOutcome) + sizeof(std::vector<double>) +
           sizeof(spectrapack::runtime::OperationControl);
}
static std::uint64_t expected_working_bytes(std::uint64_t padded_cells, std::uint64_t longest_axis,
                                            std::uint64_t reserved_bytes = 0)
{
    return reserved_bytes + metadata_bytes() + 2 * padded_cells * 16 + padded_cells * 8 + 64 * longest_axis * 16 +
           (64ULL << 10) + 8 * sizeof(std::size_t);
}

static const sc::CorrelationFailure* failed(const sc::CorrelationOutcome& o)
{
    return std::get_if<sc::CorrelationFailure>(&o);
}

// correlation_tests.cpp:184-198 binary_pattern
static std::vector<std::uint8_t> binary_pattern(sc::Shape3 shape, std::uint32_t salt)
{
    std::vector<std::uint8_t> cells(std::size_t(shape[0]) * shape[1] * shape[2]);
    for (std::uint32_t z = 0; z != shape[2]; ++z)
        for (std::uint32_t y = 0; y != shape[1]; ++y)
            for (std::uint32_t x = 0; x != shape[0]; ++x) {
                const auto hash = 17 * x + 11 * y + 5 * z + 3 * x * y + 7 * y * z + 13 * x * z + salt;
                cells[x + std::size_t(shape[0]) * (y + std::size_t(shape[1]) * z)] = std::uint8_t((hash % 19) < 7);
            }
    cells.front() = 1;
    cells.back() = 1;
    return cells;
}

// The estimate alone (no device work): estimate_correlation_cpu's working
// bytes (correlation.cpp:596-641) on the reference's own cases.
static int estimate_checks()
{
    const std::uint64_t meta = metadata_bytes();
    std::printf("metadata bytes %llu\n", (unsigned long long)meta);
    const sc::CorrelationSpec one { { 1, 1, 1 }, { 1, 1, 1 }, { 0, 0, 0 }, { 0, 0, 0 } };
    const auto e1 = sc::estimate_correlation_cpu(one, 1);
    CHECK(std::holds_alternative<sc::CorrelationEstimate>(e1));
    if (const auto* e = std::get_if<sc::CorrelationEstimate>(&e1)) {
        // correlation_tests.cpp:593: 66,664 accounted bytes plus the metadata
        CHECK(e->working_bytes == 66'664 + meta);
        CHECK(e->working_bytes == expected_working_bytes(1, 1));
        CHECK(e->padded_cells == 1);
    }
    // correlation_tests.cpp:216-239 with lengths 8 + 6 - 1 = 13 and 64 + 64 - 1 = 127 on every axis
    for (const auto [el, kl] : { std::pair<std::uint32_t, std::uint32_t> { 8, 6 }, { 64, 64 } })
        for (int axis = 0; axis != 3; ++axis) {
            sc::Shape3 es { 1, 1, 1 }, ks { 1, 1, 1 };
            es[axis] = el;
            ks[axis] = kl;
            const sc::CorrelationSpec spec { es, ks, { 6, -5, 9 }, { -3, 4, -7 } };
            const auto est = sc::estimate_correlation_cpu(spec, 1);
            const std::uint64_t p = el + kl - 1;
            CHECK(std::holds_alternative<sc::CorrelationEstimate>(est) &&
                  std::get<sc::CorrelationEstimate>(est).working_bytes == expected_working_bytes(p, p));
            CHECK(std::holds_alternative<sc::CorrelationEstimate>(est) &&
                  std::get<sc::CorrelationEstimate>(est).padded_shape[axis] == p);
        }
    // one pocketfft workspace per FFT thread (correlation.cpp:625-628)
    const sc::CorrelationSpec threads { { 5, 3, 2 }, { 2, 1, 1 }, {}, {} };
    const auto t1 = sc::estimate_correlation_cpu(threads, 1);
    for (const std::uint32_t c : { 2U, 4U, 8U }) {
        const auto tc = sc::estimate_correlation_cpu(threads, c);
        CHECK(std::holds_alternative<sc::CorrelationEstimate>(tc) &&
              std::get<sc::CorrelationEstimate>(tc).working_bytes ==
                  std::get<sc::CorrelationEstimate>(t1).working_bytes + (c - 1) * (64 * 6 * 16 + (64ULL << 10)));
    }
    for (const std::uint32_t c : { 0U, 9U }) {
        const auto bad = sc::estimate_correlation_cpu(threads, c);
        CHECK(std::holds_alternative<sc::CorrelationFailure>(bad) &&
              std::get<sc::CorrelationFailure>(bad).code == "CPU_THREAD_COUNT_UNSUPPORTED");
    }
    // representability (correlation.cpp:601-619)
    const sc::CorrelationSpec zero { { 0, 1, 1 }, { 1, 1, 1 }, {}, {} };
    const auto z = sc::estimate_correlation_cpu(zero, 1);
    CHECK(std::holds_alternative<sc::CorrelationFailure>(z) &&
          std::get<sc::CorrelationFailure>(z).code == "CORRELATION_SHAPE_OVERFLOW");
    const sc::CorrelationSpec wide { { 0xffffffffu, 1, 1 }, { 2, 1, 1 }, {}, {} };
    const auto w = sc::estimate_correlation_cpu(wide, 1);
    CHECK(std::holds_alternative<sc::CorrelationFailure>(w) &&
          std::get<sc::CorrelationFailure>(w).code == "CORRELATION_SHAPE_OVERFLOW");
    const sc::CorrelationSpec origin { { 1, 1, 1 }, { 2, 1, 1 }, { std::numeric_limits<std::int64_t>::min(), 0, 0 }, {} };
    const auto o = sc::estimate_correlation_cpu(origin, 1);
    CHECK(std::holds_alternative<sc::CorrelationFailure>(o) &&
          std::get<sc::CorrelationFailure>(o).code == "CORRELATION_INDEX_OVERFLOW");
    std::printf("estimate checks: %d failures\n", failures);
    return failures == 0 ? 0 : 1;
}

// correlation_tests.cpp:575-605 (AT-12 preflight) and :216-239
// (qualify_padded_axis): the admission a correlation applies.
static void admission_checks()
{
    const std::uint64_t meta = metadata_bytes();
    const sc::CorrelationSpec spec { { 1, 1, 1 }, { 1, 1, 1 }, { 0, 0, 0 }, { 0, 0, 0 } };
    const std::vector<std::uint8_t> one { 1 };
    sc::CorrelationLimits limits;
    limits.max_working_bytes = 1024;
    CHECK(failed(sc::correlate_binary_cpu(spec, one, one, limits)) &&
          failed(sc::correlate_binary_cpu(spec, one, one, limits))->code == "CORRELATION_MEMORY_LIMIT");
    limits.max_working_bytes = 512ULL << 20;
    limits.reserved_bytes = 1234;
    const auto ok = sc::correlate_binary_cpu(spec, one, one, limits);
    CHECK(std::holds_alternative<sc::CorrelationResult>(ok));
    const std::uint64_t accounted = 66'664 + meta;
    if (const auto* r = std::get_if<sc::CorrelationResult>(&ok))
        CHECK(r->stats.working_bytes_peak == limits.reserved_bytes + accounted);
    limits.max_working_bytes = limits.reserved_bytes + accounted - 1;
    const auto over = sc::correlate_binary_cpu(spec, one, one, limits);
    CHECK(failed(over) && failed(over)->code == "CORRELATION_MEMORY_LIMIT");
    limits.max_working_bytes = std::numeric_limits<std::uint64_t>::max();
    limits.reserved_bytes = std::numeric_limits<std::uint64_t>::max();
    const auto wrap = sc::correlate_binary_cpu(spec, one, one, limits);
    CHECK(failed(wrap) && failed(wrap)->code == "CORRELATION_MEMORY_LIMIT");
    for (const auto [el, kl] : { std::pair<std::uint32_t, std::uint32_t> { 8, 6 }, { 64, 64 } })
        for (int axis = 0; axis != 3; ++axis) {
            sc::Shape3 es { 1, 1, 1 }, ks { 1, 1, 1 };
            es[axis] = el;
            ks[axis] = kl;
            const sc::CorrelationSpec s { es, ks, { 6, -5, 9 }, { -3, 4, -7 } };
            const auto env = binary_pattern(es, 3);
            const auto ker = binary_pattern(ks, 11);
            sc::CorrelationLimits lim;
            lim.reserved_bytes = 321;
            const std::uint64_t p = el + kl - 1;
            lim.max_working_bytes = expected_working_bytes(p, p, lim.reserved_bytes);
            const auto r = sc::correlate_binary_cpu(s, env, ker, lim);
            CHECK(std::holds_alternative<sc::CorrelationResult>(r) &&
                  std::get<sc::CorrelationResult>(r).stats.working_bytes_peak == lim.max_working_bytes);
            --lim.max_working_bytes;
            const auto rej = sc::correlate_binary_cpu(s, env, ker, lim);
            CHECK(failed(rej) && failed(rej)->code == "CORRELATION_MEMORY_LIMIT");
        }
}

static bool read_scene(std::istream& in, sg::Environment& env, std::vector<sg::OrientedKernel>& kernels)
{
    for (auto& v : env.shape) in >> v;
    for (auto& v : env.first) in >> v;
    in >> env.pitch_mm;
    env.occupied.resize(std::size_t(env.shape[0]) * env.shape[1] * env.shape[2]);
    for (auto& c : env.occupied) {
        int v;
        in >> v;
        c = std::uint8_t(v);
    }
    std::size_t n;
    in >> n;
    kernels.resize(n);
    for (auto& k : kernels) {
        for (auto& v : k.rotation_xyzw) in >> v;
        for (auto& v : k.first) in >> v;
        for (auto& v : k.shape) in >> v;
        k.cells.resize(std::size_t(k.shape[0]) * k.shape[1] * k.shape[2]);
        for (auto& c : k.cells) {
            int v;
            in >> v;
            c = std::uint8_t(v);
        }
    }
    return bool(in);
}

static int rank(const char* path)
{
    std::ifstream in(path);
    sg::Environment env;
    std::vector<sg::OrientedKernel> kernels;
    if (!read_scene(in, env, kernels)) {
        std::printf("bad input\n");
        return 2;
    }
    sg::RankingRequest req;
    in >> req.lattice_origin_z_mm >> req.container_min_z_mm >> req.container_max_z_mm;
    std::vector<double> max_z(kernels.size());
    for (auto& v : max_z) in >> v;
    int reject_every;
    in >> req.capacity >> reject_every;
    if (!in) {
        std::printf("bad input\n");
        return 2;
    }
    std::uint64_t calls = 0;
    std::vector<sg::RankedPlacement> placed;
    std::vector<std::uint8_t> final_occ;
    const auto stats = sg::rank_parts(
        env, kernels, max_z, req, {},
        [&](const sg::Proposal&) {
            ++calls;
            return !(reject_every > 0 && calls % reject_every == 0);
        },
        nullptr, &placed, &final_occ);
    for (const auto& p : placed) {
        unsigned long long bits;
        std::memcpy(&bits, &p.score, sizeof bits);
        std::printf("R %u %lld %lld %lld %016llx\n", p.orientation, (long long)p.translation[0],
                    (long long)p.translation[1], (long long)p.translation[2], bits);
    }
    std::uint64_t mass = 0;
    for (auto c : final_occ) mass += c;
    std::printf("S %llu %llu %s %llu\n", (unsigned long long)stats.placed, (unsigned long long)stats.validations,
                std::string(stats.outcome).c_str(), (unsigned long long)mass);
    return 0;
}

static int place(const char* path)
{
    std::ifstream in(path);
    sg::Environment env;
    for (auto& v : env.shape) in >> v;
    for (auto& v : env.first) in >> v;
    in >> env.pitch_mm;
    env.occupied.resize(std::size_t(env.shape[0]) * env.shape[1] * env.shape[2]);
    for (auto& c : env.occupied) {
        int v;
        in >> v;
        c = std::uint8_t(v);
    }
    std::size_t n;
    in >> n;
    std::vector<sg::OrientedKernel> kernels(n);
    for (auto& k : kernels) {
        for (auto& v : k.rotation_xyzw) in >> v;
        for (auto& v : k.first) in >> v;
        for (auto& v : k.shape) in >> v;
        k.cells.resize(std::size_t(k.shape[0]) * k.shape[1] * k.shape[2]);
        for (auto& c : k.cells) {
            int v;
            in >> v;
            c = std::uint8_t(v);
        }
    }
    std::uint64_t capacity;
    int reject_every, stop_after;
    in >> capacity >> reject_every >> stop_after;
    if (!in) {
        std::printf("bad input\n");
        return 2;
    }
    int count = 0;
    std::vector<std::uint8_t> final_occ;
    const auto stats = sg::place_parts(
        env, kernels, sg::PlacementRequest { capacity, 0 }, {},
        [&](const sg::Proposal& p) {
            ++count;
            sg::GapVerdict v = sg::GapVerdict::accepted;
            if (stop_after > 0 && count >= stop_after) v = sg::GapVerdict::stop;
            else if (reject_every > 0 && count % reject_every == 0) v = sg::GapVerdict::rejected;
            std::printf("P %u %lld %lld %lld %d\n", p.orientation, (long long)p.translation[0],
                        (long long)p.translation[1], (long long)p.translation[2], int(v));
            return v;
        },
        nullptr, &final_occ);
    std::uint64_t mass = 0;
    for (auto c : final_occ) mass += c;
    std::printf("S %llu %llu %llu %s %llu\n", (unsigned long long)stats.steps, (unsigned long long)stats.proposals,
                (unsigned long long)stats.accepted, std::string(stats.outcome).c_str(), (unsigned long long)mass);
    return 0;
}

// place-exact: the scene of `place`, then capacity, the part mesh (vertex
// count, vertices, triangle count, triangles) and the box container's
// dimensions; place_parts with the exact host validation as the accept
// verdict.  Prints "P orientation tx ty tz verdict code" per proposal, the
// summary "S steps proposals accepted outcome" and "F code" for the exact
// validation of all accepted copies together.
static int place_exact(const char* path)
{
    std::ifstream in(path);
    sg::Environment env;
    std::vector<sg::OrientedKernel> kernels;
    if (!read_scene(in, env, kernels)) {
        std::printf("bad input\n");
        return 2;
    }
    std::uint64_t capacity;
    std::size_t nv, nt;
    in >> capacity >> nv;
    spectrapack::geometry::exact::Mesh part;
    part.vertices.resize(nv);
    for (auto& v : part.vertices) in >> v[0] >> v[1] >> v[2];
    in >> nt;
    part.triangles.resize(nt);
    for (auto& t : part.triangles) in >> t[0] >> t[1] >> t[2];
    spectrapack::geometry::exact::BoxContainer box;
    in >> box.dims_mm[0] >> box.dims_mm[1] >> box.dims_mm[2];
    if (!in) {
        std::printf("bad input\n");
        return 2;
    }
    spectrapack::geometry::exact::Validator validator(part, box);
    spectrapack::geometry::exact::ValidationReport last;
    const auto accept = sg::exact_accept(validator, &last);
    std::vector<spectrapack::geometry::exact::Pose> accepted;
    const auto stats = sg::place_parts(env, kernels, sg::PlacementRequest { capacity, 0 }, {},
                                       [&](const sg::Proposal& p) {
                                           const sg::GapVerdict v = accept(p);
                                           std::printf("P %u %lld %lld %lld %d %s\n", p.orientation,
                                                       (long long)p.translation[0], (long long)p.translation[1],
                                                       (long long)p.translation[2], int(v),
                                                       std::string(last.code).c_str());
                                           if (v == sg::GapVerdict::accepted)
                                               accepted.push_back({ p.translation_mm, p.rotation_xyzw });
                                           return v;
                                       });
    std::printf("S %llu %llu %llu %s\n", (unsigned long long)stats.steps, (unsigned long long)stats.proposals,
                (unsigned long long)stats.accepted, std::string(stats.outcome).c_str());
    const auto all = spectrapack::geometry::exact::validate_placements(part, accepted, box);
    std::printf("F %s %llu\n", std::string(all.code).c_str(), (unsigned long long)all.pair_tests);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc >= 3 && std::string(argv[1]) == "place-exact") return place_exact(argv[2]);
    if (argc >= 2 && std::string(argv[1]) == "correlate") return correlate_checks();
    if (argc >= 2 && std::string(argv[1]) == "estimate") return estimate_checks();
    if (argc >= 3 && std::string(argv[1]) == "place") return place(argv[2]);
    if (argc >= 3 && std::string(argv[1]) == "rank") return rank(argv[2]);
    std::printf("usage: %s correlate | place <input> | rank <input>\n", argv[0]);
    return 2;
}
