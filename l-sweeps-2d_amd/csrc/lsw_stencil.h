// Cubic-PML 5-point Helmholtz stencil at one DOF, shared by the host CSC
// assembly and the GPU assembly kernel.
//
// A = -(L1 D2x + L2 D2y) - (b1 D1x + b2 D1y) - omega^2 m        (Order 2)
//   D2 = [1,-2,1]/h^2, D1 = [-1/2,0,1/2]/h                     src/FiniteDifferences.cpp:11-122
//   L = alpha^2, b = alpha' alpha, alpha = 1/(1 + i sigma),
//   alpha' = -i sigma' / (1 + i sigma)^2                       src/PMLCubic2D.cpp:35-112
//   sigma = C ((x_int - x)/delta)^3 on the low side, C ((x - x_int)/delta)^3
//   on the high side, 0 on the closed interior                 src/PMLCubic2D.cpp:119-236
// Entry order: [0] centre, [1] (i-1,j), [2] (i+1,j), [3] (i,j-1), [4] (i,j+1).
#pragma once

#include <hip/hip_runtime.h>

namespace lsw {

struct Geo2 {
  double intDom[2][2];
  double compDom[2][2];
};

// Returns false when v lies outside the computational domain
// (the reference aborts: "Evaluation point out of range", PMLCubic2D.cpp:142-146).
__host__ __device__ inline bool pmlSigma(double v, double loC, double loI, double hiI, double hiC,
                                         double C, double& s, double& ds) {
  if (v >= loC && v < loI) {
    const double d = loI - loC;
    const double r = (loI - v) / d;
    s = C * r * r * r;
    ds = -3.0 * C * (loI - v) * (loI - v) / (d * d * d);
    return true;
  }
  if (v >= loI && v <= hiI) {
    s = 0.0;
    ds = 0.0;
    return true;
  }
  if (v > hiI && v < hiC) {
    const double d = hiC - hiI;
    const double r = (v - hiI) / d;
    s = C * r * r * r;
    ds = 3.0 * C * (v - hiI) * (v - hiI) / (d * d * d);
    return true;
  }
  s = 0.0;
  ds = 0.0;
  return false;
}

// alpha^2 and alpha'*alpha for one axis (re, im pairs).
__host__ __device__ inline void pmlAxis(double s, double ds, double& Lr, double& Li, double& Br,
                                        double& Bi) {
  // alpha = 1/(1 + i s) = (1 - i s)/(1 + s^2)
  const double den = 1.0 + s * s;
  const double ar = 1.0 / den, ai = -s / den;
  // alpha^2
  Lr = ar * ar - ai * ai;
  Li = 2.0 * ar * ai;
  // (1 + i s)^2 = (1 - s^2) + 2 i s ; alpha' = -i ds / (1 + i s)^2 = -i ds * alpha^2
  const double dr = ds * Li, di = -ds * Lr;  // -i*ds*(Lr + i Li) = ds*Li - i ds*Lr
  // b = alpha' * alpha
  Br = dr * ar - di * ai;
  Bi = dr * ai + di * ar;
}

// Whether every DOF of a grid (i, j) at (x0 + (i+1) hx, y0 + (j+1) hy),
// 0 <= i < nx, 0 <= j < ny, lies inside the computational domain, the
// condition under which helmholtzStencil evaluates the PML; it is separable,
// so a host check of the nx + ny coordinates stands for the nx * ny DOFs.
inline bool gridInsidePml(const Geo2& g, double x0, double y0, double hx, double hy, int nx, int ny) {
  double s, ds;
  for (int i = 0; i < nx; ++i)
    if (!pmlSigma(x0 + (double)(i + 1) * hx, g.compDom[0][0], g.intDom[0][0], g.intDom[0][1], g.compDom[0][1], 1.0,
                  s, ds))
      return false;
  for (int j = 0; j < ny; ++j)
    if (!pmlSigma(y0 + (double)(j + 1) * hy, g.compDom[1][0], g.intDom[1][0], g.intDom[1][1], g.compDom[1][1], 1.0,
                  s, ds))
      return false;
  return true;
}

// coef: 5 complex entries as (re, im) pairs.  Returns false if (x, y) is
// outside the computational domain.
__host__ __device__ inline bool helmholtzStencil(double x, double y, const Geo2& g, double C,
                                                 double omega, double m, double hx, double hy,
                                                 double coef[10]) {
  double sx, dsx, sy, dsy;
  const bool okx = pmlSigma(x, g.compDom[0][0], g.intDom[0][0], g.intDom[0][1], g.compDom[0][1],
                            C, sx, dsx);
  const bool oky = pmlSigma(y, g.compDom[1][0], g.intDom[1][0], g.intDom[1][1], g.compDom[1][1],
                            C, sy, dsy);
  double L1r, L1i, B1r, B1i, L2r, L2i, B2r, B2i;
  pmlAxis(sx, dsx, L1r, L1i, B1r, B1i);
  pmlAxis(sy, dsy, L2r, L2i, B2r, B2i);
  const double ax = 1.0 / (hx * hx), ay = 1.0 / (hy * hy);
  const double bx = 0.5 / hx, by = 0.5 / hy;
  // centre: -(L1 (-2/hx^2)) - L2 (-2/hy^2) - omega^2 m
  coef[0] = 2.0 * ax * L1r + 2.0 * ay * L2r - omega * omega * m;
  coef[1] = 2.0 * ax * L1i + 2.0 * ay * L2i;
  // (i-1): -L1/hx^2 + b1/(2hx) ; (i+1): -L1/hx^2 - b1/(2hx)
  coef[2] = -ax * L1r + bx * B1r;
  coef[3] = -ax * L1i + bx * B1i;
  coef[4] = -ax * L1r - bx * B1r;
  coef[5] = -ax * L1i - bx * B1i;
  coef[6] = -ay * L2r + by * B2r;
  coef[7] = -ay * L2i + by * B2i;
  coef[8] = -ay * L2r - by * B2r;
  coef[9] = -ay * L2i - by * B2i;
  return okx && oky;
}

}  // namespace lsw
