// Test driver of the C++ host mirror (run by tests/test_gpu_host.py).
//   spectrapack_host_test correlate          — golden vectors and failure variants
//   spectrapack_host_test place <input.txt>  — the placement loop on a scene
//     read from whitespace-separated numbers; prints one line per proposal
//     "P orientation tx ty tz verdict" and a final "S steps proposals accepted outcome".
//   spectrapack_host_test rank <input.txt>   — the ranked loop on a scene (the
//     same numbers, then origin_z cmin_z cmax_z, one max z per orientation,
//     capacity and reject_every); prints "R orientation tx ty tz score_bits"
//     per placement and a final "S placed validations outcome mass".
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
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
    std::printf("correlate checks: %d failures\n", failures);
    return failures == 0 ? 0 : 1;
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

int main(int argc, char** argv)
{
    if (argc >= 2 && std::string(argv[1]) == "correlate") return correlate_checks();
    if (argc >= 3 && std::string(argv[1]) == "place") return place(argv[2]);
    if (argc >= 3 && std::string(argv[1]) == "rank") return rank(argv[2]);
    std::printf("usage: %s correlate | place <input> | rank <input>\n", argv[0]);
    return 2;
}
