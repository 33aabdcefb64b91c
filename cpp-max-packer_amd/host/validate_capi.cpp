// C ABI over spectrapack::geometry::exact (include/spectrapack_validate.h).
#include <cstring>
#include <memory>
#include <new>

#include "spectrapack/geometry/exact_validation.hpp"
#include "spectrapack_validate.h"

namespace ex = spectrapack::geometry::exact;

namespace {

bool read_mesh(const double* v, std::uint32_t nv, const std::uint32_t* t, std::uint32_t nt, ex::Mesh& m)
{
    if ((nv && !v) || (nt && !t)) return false;
    m.vertices.resize(nv);
    for (std::uint32_t i = 0; i != nv; ++i) m.vertices[i] = { v[3 * i], v[3 * i + 1], v[3 * i + 2] };
    m.triangles.resize(nt);
    for (std::uint32_t i = 0; i != nt; ++i) {
        m.triangles[i] = { t[3 * i], t[3 * i + 1], t[3 * i + 2] };
        for (std::uint32_t k : m.triangles[i])
            if (k >= nv) return false;
    }
    return true;
}

void report(const ex::ValidationReport& r, sp_validation_report* out)
{
    *out = sp_validation_report {};
    out->validity = r.validity == ex::Validity::valid ? SP_VALID : r.validity == ex::Validity::invalid ? SP_INVALID
                                                                                                     : SP_INDETERMINATE;
    out->first = r.first == ex::kNoCopy ? UINT64_MAX : r.first;
    out->second = r.second == ex::kNoCopy ? UINT64_MAX : r.second;
    out->pair_tests = r.pair_tests;
    out->triangle_tests = r.triangle_tests;
    out->exact_evaluations = r.exact_evaluations;
    std::strncpy(out->code, std::string(r.code).c_str(), sizeof out->code - 1);
}

ex::Pose pose_of(const sp_pose& p)
{
    ex::Pose q;
    for (int a = 0; a != 3; ++a) q.translation_mm[a] = p.translation_mm[a];
    for (int a = 0; a != 4; ++a) q.rotation_xyzw[a] = p.rotation_xyzw[a];
    return q;
}

}  // namespace

struct sp_validator {
    ex::Mesh part, container;
    std::unique_ptr<ex::Validator> v;
};

extern "C" {

int sp_validate_placements(const double* pv, uint32_t npv, const uint32_t* pt, uint32_t npt, const sp_pose* poses,
                           uint64_t n, const double* box, const double* cv, uint32_t ncv, const uint32_t* ct,
                           uint32_t nct, double pair_clearance, double wall_clearance, sp_validation_report* out)
{
    if (!out || (n && !poses)) return -1;
    try {
        ex::Mesh part, cont;
        if (!read_mesh(pv, npv, pt, npt, part)) return -1;
        ex::Container c = ex::BoxContainer {};
        if (box) c = ex::BoxContainer { { box[0], box[1], box[2] } };
        else {
            if (!read_mesh(cv, ncv, ct, nct, cont) || nct == 0) return -1;
            c = &cont;
        }
        std::vector<ex::Pose> ps(n);
        for (uint64_t i = 0; i != n; ++i) ps[i] = pose_of(poses[i]);
        report(ex::validate_placements(part, ps, c, pair_clearance, wall_clearance), out);
        return 0;
    }
    catch (const std::bad_alloc&) {
        return -1;
    }
}

int sp_validator_create(const double* pv, uint32_t npv, const uint32_t* pt, uint32_t npt, const double* box,
                        const double* cv, uint32_t ncv, const uint32_t* ct, uint32_t nct, double pair_clearance,
                        double wall_clearance, sp_validator** out)
{
    if (!out) return -1;
    *out = nullptr;
    try {
        auto s = std::make_unique<sp_validator>();
        if (!read_mesh(pv, npv, pt, npt, s->part)) return -1;
        ex::Container c = ex::BoxContainer {};
        if (box) c = ex::BoxContainer { { box[0], box[1], box[2] } };
        else {
            if (!read_mesh(cv, ncv, ct, nct, s->container) || nct == 0) return -1;
            c = &s->container;
        }
        s->v = std::make_unique<ex::Validator>(s->part, c, pair_clearance, wall_clearance);
        *out = s.release();
        return 0;
    }
    catch (const std::bad_alloc&) {
        return -1;
    }
}

int sp_validator_try_add(sp_validator* v, const sp_pose* pose, sp_validation_report* out)
{
    if (!v || !pose || !out) return -1;
    try {
        report(v->v->try_add(pose_of(*pose)), out);
        return 0;
    }
    catch (const std::bad_alloc&) {
        return -1;
    }
}

void sp_validator_destroy(sp_validator* v) { delete v; }

}  // extern "C"
