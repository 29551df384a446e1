// g2o dataset I/O with gtsam::load3D / load2D / writeG2o semantics (the
// reference reads its fixtures with gtsam::load3D, tests/testGnc.cpp:31, and
// writes with src/Logger.cpp:64-163):
//   * VERTEX_SE3:QUAT id x y z qx qy qz qw, EDGE_SE3:QUAT i j x y z qx qy qz qw
//     + 21 upper-triangular information entries in g2o's (t, R) order,
//     reordered to the (R, t) tangent order;
//   * VERTEX_SE2 id x y theta, EDGE_SE2 i j x y theta + 6 information entries
//     in (x, y, theta) order.
// Keys are the file's integers (gtsam::Symbol packing).
#pragma once

#include <string>
#include <utility>

#include "kimera_rpgo_amd/Types.h"

namespace kimera_rpgo_amd {

using GraphAndValues = std::pair<NonlinearFactorGraph, Values>;

GraphAndValues load3D(const std::string& path);
GraphAndValues load2D(const std::string& path);
void writeG2o(const NonlinearFactorGraph& graph, const Values& values, const std::string& path);

}  // namespace kimera_rpgo_amd
