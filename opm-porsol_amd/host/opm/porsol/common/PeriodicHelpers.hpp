// Periodic boundary set-up with the behaviour of
// opm/porsol/common/PeriodicHelpers.hpp (createPeriodic :105-181,
// createPeriodicImpl :183-214, storeFlowCond :49-64) and
// opm/porsol/common/BoundaryPeriodicity.{hpp,cpp} (findPeriodicPartners
// :86-175, match :25-48): boundary faces are classified by the side of the
// centroid bounding box they lie on, sorted by a key built from their two
// in-plane centroid coordinates, and paired with the face on the opposite side
// whose area and in-plane centroid agree within the tolerance (first a window
// of +-10 sorted positions, then a full scan).  Requires unique boundary ids.
#ifndef OPM_PORSOL_AMD_PERIODICHELPERS_HPP
#define OPM_PORSOL_AMD_PERIODICHELPERS_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "BoundaryConditions.hpp"

namespace Opm {

struct BoundaryFaceInfo {
    int face_index;          // position in the boundary-face list
    int bid;                 // boundary id
    int canon_pos;           // 0 xmin, 1 xmax, 2 ymin, ...
    int partner_face_index;  // -1 if none
    int partner_bid;         // 0 if none
    double area;
    double centroid[3];      // with the coordinate normal to its side zeroed

    bool operator<(const BoundaryFaceInfo &o) const { return cmpval() < o.cmpval(); }

private:
    double cmpval() const {
        const double pi = 3.14159265358979323846264338327950288;
        return centroid[(canon_pos / 2 + 1) % 3] + pi * centroid[(canon_pos / 2 + 2) % 3];
    }
};

inline bool matchPeriodic(std::vector<BoundaryFaceInfo> &bf, int face, int lower, int upper) {
    const double area_tol = 1e-6, centroid_tol = 1e-6;
    const int cp = bf[face].canon_pos;
    const int target = (cp % 2 == 0) ? cp + 1 : cp - 1;
    for (int j = lower; j < upper; ++j) {
        if (bf[j].canon_pos != target) continue;
        if (std::fabs(bf[face].area - bf[j].area) > area_tol) continue;
        double d2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double t = bf[j].centroid[d] - bf[face].centroid[d];
            d2 += t * t;
        }
        if (std::sqrt(d2) <= centroid_tol) {
            bf[face].partner_face_index = bf[j].face_index;
            bf[face].partner_bid = bf[j].bid;
            bf[j].partner_face_index = bf[face].face_index;
            bf[j].partner_bid = bf[face].bid;
            break;
        }
    }
    return bf[face].partner_face_index != -1;
}

template <class GridInterface>
void findPeriodicPartners(std::vector<BoundaryFaceInfo> &bfinfo, std::array<double, 6> &side_areas,
                          const GridInterface &g, const std::array<bool, 6> &is_periodic,
                          double spatial_tolerance = 1e-6) {
    typedef typename GridInterface::CellIterator CI;
    typedef typename CI::FaceIterator FI;
    struct Raw { int bid; double area; double c[3]; };
    std::vector<Raw> raw;
    double lo[3] = {1e100, 1e100, 1e100}, hi[3] = {-1e100, -1e100, -1e100};
    int max_bid = 0;
    for (CI c = g.cellbegin(); c != g.cellend(); ++c)
        for (FI f = c->facebegin(); f != c->faceend(); ++f) {
            if (!f->boundaryId()) continue;
            const auto x = f->centroid();
            Raw r{f->boundaryId(), f->area(), {x[0], x[1], x[2]}};
            raw.push_back(r);
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], x[d]);
                hi[d] = std::max(hi[d], x[d]);
            }
            max_bid = std::max(max_bid, f->boundaryId());
        }
    const int num_bdy = int(raw.size());
    if (max_bid != num_bdy)
        throw std::runtime_error("createPeriodic() assumes that every boundary face has a unique boundary id. That seems to be violated.");
    side_areas.fill(0.0);
    bfinfo.clear();
    for (int i = 0; i < num_bdy; ++i) {
        BoundaryFaceInfo bf;
        bf.face_index = i;
        bf.bid = raw[i].bid;
        bf.canon_pos = -1;
        bf.partner_face_index = -1;
        bf.partner_bid = 0;
        bf.area = raw[i].area;
        for (int d = 0; d < 3; ++d) bf.centroid[d] = raw[i].c[d];
        for (int d = 0; d < 3; ++d) {
            if (std::fabs(bf.centroid[d] - lo[d]) <= spatial_tolerance) { bf.canon_pos = 2 * d; break; }
            if (std::fabs(bf.centroid[d] - hi[d]) <= spatial_tolerance) { bf.canon_pos = 2 * d + 1; break; }
        }
        if (bf.canon_pos == -1)
            throw std::runtime_error("Boundary face centroid not on bounding box. Maybe the grid is not an axis-aligned shoe-box?");
        side_areas[bf.canon_pos] += bf.area;
        bf.centroid[bf.canon_pos / 2] = 0.0;
        bfinfo.push_back(bf);
    }
    std::sort(bfinfo.begin(), bfinfo.end());
    for (int i = 0; i < num_bdy; ++i) {
        if (bfinfo[i].partner_face_index != -1) continue;
        if (!is_periodic[bfinfo[i].canon_pos]) continue;
        const int lower = std::max(0, i - 10), upper = std::min(num_bdy, i + 10);
        if (!matchPeriodic(bfinfo, i, lower, upper) && !matchPeriodic(bfinfo, i, 0, num_bdy))
            std::cerr << "Warning: No partner found for boundary id " << bfinfo[i].bid << '\n';
    }
}

/// Periodic flow conditions on the sides marked periodic in fconditions
/// (createPeriodic, PeriodicHelpers.hpp:135-157).  Neumann values are scaled by
/// face area / side area (storeFlowCond :49-64).
template <class BCs, class GridInterface>
void createPeriodic(BCs &fbcs, const GridInterface &g, const std::array<FlowBC, 6> &fconditions,
                    double spatial_tolerance = 1e-6) {
    std::array<bool, 6> is_periodic;
    for (int i = 0; i < 6; ++i) is_periodic[i] = fconditions[i].isPeriodic();
    for (int i = 0; i < 3; ++i)
        if (fconditions[2 * i].isPeriodic() &&
            (!fconditions[2 * i + 1].isPeriodic() ||
             fconditions[2 * i].pressureDifference() != -fconditions[2 * i + 1].pressureDifference()))
            throw std::runtime_error("createPeriodic: opposite sides need periodic conditions with opposite pressure differences");
    std::vector<BoundaryFaceInfo> bfinfo;
    std::array<double, 6> side_areas;
    findPeriodicPartners(bfinfo, side_areas, g, is_periodic, spatial_tolerance);
    const int num_bdy = int(bfinfo.size());
    fbcs.clear();
    fbcs.resize(num_bdy + 1);
    for (int i = 0; i < num_bdy; ++i) {
        if (bfinfo[i].bid < bfinfo[i].partner_bid) fbcs.setPeriodicPartners(bfinfo[i].bid, bfinfo[i].partner_bid);
        fbcs.setCanonicalBoundaryId(bfinfo[i].bid, bfinfo[i].canon_pos + 1);
    }
    for (int i = 0; i < num_bdy; ++i) {
        FlowBC face_bc = fconditions[bfinfo[i].canon_pos];
        if (face_bc.isNeumann())
            face_bc = FlowBC(FlowBC::Neumann, face_bc.outflux() * bfinfo[i].area / side_areas[bfinfo[i].canon_pos]);
        fbcs.flowCond(bfinfo[i].bid) = face_bc;
    }
}

}  // namespace Opm

#endif
