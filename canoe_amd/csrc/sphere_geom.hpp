// sphere_geom.hpp -- host-side geometry of the cubed-sphere run (config 3):
// per-panel metric tables, the ghost-zone interpolation table and the
// panel-edge exchange table that the device kernels of sphere.hpp read.
//
// Panel p = 1..6 covers x2, x3 in [-pi/4, pi/4] (gnomonic equiangle,
// src/exo3/gnomonic_equiangle.cpp:24-219); its points are those of the
// reference's GetLatLon (src/exo3/cubed_sphere.cpp:78-89, RLLFromXYP of
// cubed_sphere_transform.cpp:59-118) with X = tan(x2), Y = tan(x3).  A point beyond a panel's
// edge is located on the neighbour by inverting the neighbour's GetLatLon,
// and vectors cross between panels through the reference's (U, V)
// components (VecTransRLLFromABP / VecTransABPFromRLL,
// cubed_sphere_transform.cpp:11-57) at that point.
#pragma once
#include <vector>

namespace cg_host {

// (k, j) tables of one panel, each nc x nc (nc = n + 2 NGHOST), index
// k * nc + j; the order of the device's kj array.
enum SphKJ {
  SKJ_COSC = 0, SKJ_SINC, SKJ_COSF2, SKJ_SINF2, SKJ_COSF3, SKJ_SINF3,
  SKJ_D2A, SKJ_D3A, SKJ_D2AF3, SKJ_D3AF2,
  SKJ_HC, SKJ_HF2, SKJ_HF3,  // cos(theta) of hllc_transform.cpp's projection
  SKJ_SINLAT, SKJ_COSLAT,
  SKJ_R00, SKJ_R01, SKJ_R10, SKJ_R11,  // (v2, v3) -> (U, V), VecTransRLLFromABP
  SKJ_A00, SKJ_A01, SKJ_A10, SKJ_A11,  // (U, V) -> (v2, v3), VecTransABPFromRLL
  SKJ_COUNT
};

// One ghost cell column of a panel (all interior x1 levels): the
// interpolation along row `row` of the neighbour q between cells a1 and a2
// (y: the row angle of CalculateInterpLocations at a1, a2 and the ghost),
// and the covariant momentum map from q's frame to p's at the ghost point.
struct SphGhost {
  int p, kg, jg;          // panel (0-based) and the ghost's (k, j)
  int q, k1, j1, k2, j2;  // neighbour panel (0-based) and the two cells
  double y1, y2, yg;
  double sq, cq;          // sin, cos(theta) of q's frame at the ghost point
  double r00, r01, r10, r11;  // VecTransRLLFromABP there (panel q)
  double a00, a01, a10, a11;  // VecTransABPFromRLL at the ghost in p's frame
  double cp;              // cos(theta) of p's frame at the ghost point
};

// One panel-edge face column: the outer face state of panel p at its side s
// (0: x2 lower, 1: x2 upper, 2: x3 lower, 3: x3 upper) and along index t
// takes q's inner state at its side sq and index tq, the contravariant
// velocity carried from q's frame to p's through (U, V) at the face.
struct SphEdge {
  int p, s, t, q, sq, tq;
  double r00, r01, r10, r11;  // VecTransRLLFromABP at the face in q's frame
  double a00, a01, a10, a11;  // VecTransABPFromRLL at the face in p's frame
};

struct SphGeom {
  int n, nc, nr, nc1;
  std::vector<double> kj[6];  // SKJ_COUNT * nc * nc per panel
  std::vector<double> x1f, x1v;
  std::vector<SphGhost> ghosts;
  std::vector<SphEdge> edges;
  int link_q[6][4], link_sq[6][4], link_inv[6][4];
};

// n cells per panel side, nr levels from radius r0 with spacing dx1, omega
// for the Coriolis tables.
SphGeom sphere_geometry(int n, int nr, double r0, double dx1);

}  // namespace cg_host
