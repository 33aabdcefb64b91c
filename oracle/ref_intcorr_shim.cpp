// C entry point over the reference's direct integer correlation oracle
// (tests/support/include/spectrapack/test_support/integer_correlation.hpp:76-79),
// so tests can compare it with the oracle's restatement.  Built only by
// `make ref`, into oracle/_ref/.
#include <cstring>
#include <stdexcept>

#include "spectrapack/test_support/integer_correlation.hpp"

namespace ts = spectrapack::test_support;

extern "C" int ref_direct_correlation(const int* b, const int* be, const int* k, const int* ke, const int* origin,
                                      int* range, int* extent, int* values, unsigned long long capacity)
{
    try {
        const ts::Extent3 bx { be[0], be[1], be[2] }, kx { ke[0], ke[1], ke[2] };
        const ts::IntField3 blocked { bx, std::vector<int>(b, b + bx.cell_count()) };
        const ts::IntField3 kernel { kx, std::vector<int>(k, k + kx.cell_count()) };
        const auto r = ts::direct_linear_cross_correlation(blocked, kernel, { origin[0], origin[1], origin[2] });
        const int rr[6] = { r.translations.min_x, r.translations.max_x, r.translations.min_y,
                            r.translations.max_y, r.translations.min_z, r.translations.max_z };
        std::memcpy(range, rr, sizeof rr);
        extent[0] = r.extent.x;
        extent[1] = r.extent.y;
        extent[2] = r.extent.z;
        if (r.values.size() > capacity) return 3;
        std::memcpy(values, r.values.data(), r.values.size() * sizeof(int));
        return 0;
    }
    catch (const std::invalid_argument&) {
        return 1;
    }
    catch (const std::overflow_error&) {
        return 2;
    }
}
