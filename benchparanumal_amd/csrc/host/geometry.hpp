// Geometric factors at one point from the 9 coordinate derivatives, shared
// by the host setup (host/mesh.cpp) and the device setup (setup.hip) so that
// both evaluate the same expressions in the same order; with floating-point
// contraction off in both builds the two give the same bits.
//
// Reference: libs/mesh/meshGeometricFactorsHex3D.cpp:116-135 (J by cofactor
// expansion, the inverse Jacobian's rows by cofactors over J, and the six
// entries of J (grad r)(grad r)^T); the formula is the textbook one.
#pragma once

#if defined(__HIPCC__)
#define BPX_HD __host__ __device__
#else
#define BPX_HD
#endif

namespace bpx {

struct Geo9 {
  double xr, xs, xt, yr, ys, yt, zr, zs, zt;
};

// J and the 6 symmetric entries of J (grad r)(grad r)^T
BPX_HD inline void metric(const Geo9& d, double& J, double g[6]) {
  J = d.xr * (d.ys * d.zt - d.zs * d.yt) - d.yr * (d.xs * d.zt - d.zs * d.xt) +
      d.zr * (d.xs * d.yt - d.ys * d.xt);
  const double rx = (d.ys * d.zt - d.zs * d.yt) / J, ry = -(d.xs * d.zt - d.zs * d.xt) / J,
               rz = (d.xs * d.yt - d.ys * d.xt) / J;
  const double sx = -(d.yr * d.zt - d.zr * d.yt) / J, sy = (d.xr * d.zt - d.zr * d.xt) / J,
               sz = -(d.xr * d.yt - d.yr * d.xt) / J;
  const double tx = (d.yr * d.zs - d.zr * d.ys) / J, ty = -(d.xr * d.zs - d.zr * d.xs) / J,
               tz = (d.xr * d.ys - d.yr * d.xs) / J;
  g[0] = J * (rx * rx + ry * ry + rz * rz);
  g[1] = J * (rx * sx + ry * sy + rz * sz);
  g[2] = J * (rx * tx + ry * ty + rz * tz);
  g[3] = J * (sx * sx + sy * sy + sz * sz);
  g[4] = J * (sx * tx + sy * ty + sz * tz);
  g[5] = J * (tx * tx + ty * ty + tz * tz);
}

// sign (+1 / -1) of the reference coordinates r, s, t of vertex v (0..7) in
// the reference's vertex order (libs/mesh/meshSetElementType.cpp)
BPX_HD constexpr double vertex_r(int v) { return ((v ^ (v >> 1)) & 1) ? 1.0 : -1.0; }
BPX_HD constexpr double vertex_s(int v) { return ((v >> 1) & 1) ? 1.0 : -1.0; }
BPX_HD constexpr double vertex_t(int v) { return ((v >> 2) & 1) ? 1.0 : -1.0; }

// the trilinear map's weight of vertex v at (r, s, t)
BPX_HD inline double vertex_weight(int v, double r, double s, double t) {
  return 0.125 * (1 + vertex_r(v) * r) * (1 + vertex_s(v) * s) * (1 + vertex_t(v) * t);
}

}  // namespace bpx
