// g2o reader / writer (see Dataset.h for the formats).
#include "kimera_rpgo_amd/Dataset.h"

#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace kimera_rpgo_amd {

static void read_upper(std::istringstream& is, int d, std::vector<double>& m, const std::string& path) {
  m.assign(d * d, 0.0);
  for (int a = 0; a < d; ++a)
    for (int b = a; b < d; ++b) {
      double v;
      if (!(is >> v)) throw std::runtime_error(path + ": truncated information matrix");
      m[a * d + b] = v;
      m[b * d + a] = v;
    }
}

GraphAndValues load3D(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  GraphAndValues out;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string tag;
    if (!(is >> tag)) continue;
    if (tag == "VERTEX_SE3:QUAT") {
      unsigned long long id;
      double x, y, z, qx, qy, qz, qw;
      if (!(is >> id >> x >> y >> z >> qx >> qy >> qz >> qw)) throw std::runtime_error(path + ": bad vertex");
      out.second.insert((Key)id, Pose3::FromQuaternion(qw, qx, qy, qz, x, y, z));
    } else if (tag == "EDGE_SE3:QUAT") {
      unsigned long long i, j;
      double x, y, z, qx, qy, qz, qw;
      if (!(is >> i >> j >> x >> y >> z >> qx >> qy >> qz >> qw)) throw std::runtime_error(path + ": bad edge");
      std::vector<double> m;
      read_upper(is, 6, m, path);
      // g2o (t, R) -> tangent (R, t): swap the 3x3 blocks
      std::vector<double> info(36);
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) info[((a + 3) % 6) * 6 + (b + 3) % 6] = m[a * 6 + b];
      out.first.add(BetweenFactorPose3((Key)i, (Key)j, Pose3::FromQuaternion(qw, qx, qy, qz, x, y, z),
                                       NoiseModel::Information(6, info)));
    }
  }
  return out;
}

GraphAndValues load2D(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  GraphAndValues out;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string tag;
    if (!(is >> tag)) continue;
    if (tag == "VERTEX_SE2") {
      unsigned long long id;
      double x, y, th;
      if (!(is >> id >> x >> y >> th)) throw std::runtime_error(path + ": bad vertex");
      out.second.insert((Key)id, Pose2(x, y, th));
    } else if (tag == "EDGE_SE2") {
      unsigned long long i, j;
      double x, y, th;
      if (!(is >> i >> j >> x >> y >> th)) throw std::runtime_error(path + ": bad edge");
      std::vector<double> m;
      read_upper(is, 3, m, path);
      out.first.add(BetweenFactorPose2((Key)i, (Key)j, Pose2(x, y, th), NoiseModel::Information(3, m)));
    }
  }
  return out;
}

// rotation matrix -> unit quaternion (w, x, y, z), Shepperd's method
static void rot_to_quat(const std::array<double, 9>& R, double q[4]) {
  const double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    const double s = 0.5 / std::sqrt(tr + 1.0);
    q[0] = 0.25 / s;
    q[1] = (R[7] - R[5]) * s;
    q[2] = (R[2] - R[6]) * s;
    q[3] = (R[3] - R[1]) * s;
    return;
  }
  int i = 0;
  if (R[4] > R[0]) i = 1;
  if (R[8] > R[i * 4]) i = 2;
  const int j = (i + 1) % 3, k = (i + 2) % 3;
  const double s = 2.0 * std::sqrt(1.0 + R[i * 4] - R[j * 4] - R[k * 4]);
  double v[3];
  v[i] = 0.25 * s;
  v[j] = (R[j * 3 + i] + R[i * 3 + j]) / s;
  v[k] = (R[k * 3 + i] + R[i * 3 + k]) / s;
  q[0] = (R[k * 3 + j] - R[j * 3 + k]) / s;
  q[1] = v[0];
  q[2] = v[1];
  q[3] = v[2];
}

void writeG2o(const NonlinearFactorGraph& graph, const Values& values, const std::string& path) {
  std::ofstream f(path);
  if (!f) throw std::runtime_error("cannot write " + path);
  f << std::setprecision(17);
  for (const auto& kv : values.pose3s()) {
    double q[4];
    rot_to_quat(kv.second.R, q);
    f << "VERTEX_SE3:QUAT " << kv.first << ' ' << kv.second.t[0] << ' ' << kv.second.t[1] << ' '
      << kv.second.t[2] << ' ' << q[1] << ' ' << q[2] << ' ' << q[3] << ' ' << q[0] << '\n';
  }
  for (const auto& kv : values.pose2s())
    f << "VERTEX_SE2 " << kv.first << ' ' << kv.second.x << ' ' << kv.second.y << ' ' << kv.second.theta << '\n';
  for (const auto& fp : graph) {
    const Factor& fa = *fp;
    if (fa.type == FactorType::BETWEEN_POSE3) {
      double q[4];
      rot_to_quat(fa.pose3.R, q);
      f << "EDGE_SE3:QUAT " << fa.key1 << ' ' << fa.key2 << ' ' << fa.pose3.t[0] << ' ' << fa.pose3.t[1] << ' '
        << fa.pose3.t[2] << ' ' << q[1] << ' ' << q[2] << ' ' << q[3] << ' ' << q[0];
      const auto& I = fa.noise->information();
      for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) f << ' ' << I[((a + 3) % 6) * 6 + (b + 3) % 6];
      f << '\n';
    } else if (fa.type == FactorType::BETWEEN_POSE2) {
      f << "EDGE_SE2 " << fa.key1 << ' ' << fa.key2 << ' ' << fa.pose2.x << ' ' << fa.pose2.y << ' '
        << fa.pose2.theta;
      const auto& I = fa.noise->information();
      for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) f << ' ' << I[a * 3 + b];
      f << '\n';
    }
  }
}

}  // namespace kimera_rpgo_amd
