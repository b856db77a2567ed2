// delfem_amd/ker_emat.hpp -- host versions of DelFEM's P1 element helpers,
// for element loops written against the reference (AddLinSys_* in femeqn):
//
//   TriArea, TriDlDx       src/femeqn/ker_emat_tri.cpp:12-14, 36-69
//   TetVolume, TetDlDx     include/delfem/femeqn/ker_emat_tet.h:32-77
//
// The arithmetic follows the reference term by term (the device assembly
// kernels use the same expressions), so element matrices computed with these
// match the reference's bit for bit when compiled without FMA contraction
// (-ffp-contract=off, the default of g++ on x86-64 without -march flags).
#pragma once

// signed area of the triangle (p0, p1, p2)
inline double TriArea(const double p0[], const double p1[], const double p2[]) {
  return 0.5 * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}

// gradients of the three area coordinates and their constant terms
inline void TriDlDx(double dldx[][2], double const_term[], const double p0[], const double p1[], const double p2[]) {
  const double t = 0.5 / TriArea(p0, p1, p2);
  const_term[0] = t * (p1[0] * p2[1] - p2[0] * p1[1]);
  const_term[1] = t * (p2[0] * p0[1] - p0[0] * p2[1]);
  const_term[2] = t * (p0[0] * p1[1] - p1[0] * p0[1]);
  dldx[0][0] = t * (p1[1] - p2[1]);
  dldx[1][0] = t * (p2[1] - p0[1]);
  dldx[2][0] = t * (p0[1] - p1[1]);
  dldx[0][1] = t * (p2[0] - p1[0]);
  dldx[1][1] = t * (p0[0] - p2[0]);
  dldx[2][1] = t * (p1[0] - p0[0]);
}

// signed volume of the tetrahedron (p0, p1, p2, p3)
inline double TetVolume(const double p0[], const double p1[], const double p2[], const double p3[]) {
  double v = (p1[0] - p0[0]) * ((p2[1] - p0[1]) * (p3[2] - p0[2]) - (p3[1] - p0[1]) * (p2[2] - p0[2])) -
             (p1[1] - p0[1]) * ((p2[0] - p0[0]) * (p3[2] - p0[2]) - (p3[0] - p0[0]) * (p2[2] - p0[2])) +
             (p1[2] - p0[2]) * ((p2[0] - p0[0]) * (p3[1] - p0[1]) - (p3[0] - p0[0]) * (p2[1] - p0[1]));
  v *= 0.1666666666666;
  return v;
}

// gradients of the four volume coordinates
inline void TetDlDx(double dldx[][3], double a[], const double p0[], const double p1[], const double p2[],
                    const double p3[]) {
  const double vol = TetVolume(p0, p1, p2, p3);
  const double t = 1.0 / (vol * 6.0);
  a[0] = +t * (p1[0] * (p2[1] * p3[2] - p3[1] * p2[2]) - p1[1] * (p2[0] * p3[2] - p3[0] * p2[2]) +
               p1[2] * (p2[0] * p3[1] - p3[0] * p2[1]));
  a[1] = -t * (p2[0] * (p3[1] * p0[2] - p0[1] * p3[2]) - p2[1] * (p3[0] * p0[2] - p0[0] * p3[2]) +
               p2[2] * (p3[0] * p0[1] - p0[0] * p3[1]));
  a[2] = +t * (p3[0] * (p0[1] * p1[2] - p1[1] * p0[2]) - p3[1] * (p0[0] * p1[2] - p1[0] * p0[2]) +
               p3[2] * (p0[0] * p1[1] - p1[0] * p0[1]));
  a[3] = -t * (p0[0] * (p1[1] * p2[2] - p2[1] * p1[2]) - p0[1] * (p1[0] * p2[2] - p2[0] * p1[2]) +
               p0[2] * (p1[0] * p2[1] - p2[0] * p1[1]));
  dldx[0][0] = -t * ((p2[1] - p1[1]) * (p3[2] - p1[2]) - (p3[1] - p1[1]) * (p2[2] - p1[2]));
  dldx[0][1] = +t * ((p2[0] - p1[0]) * (p3[2] - p1[2]) - (p3[0] - p1[0]) * (p2[2] - p1[2]));
  dldx[0][2] = -t * ((p2[0] - p1[0]) * (p3[1] - p1[1]) - (p3[0] - p1[0]) * (p2[1] - p1[1]));
  dldx[1][0] = +t * ((p3[1] - p2[1]) * (p0[2] - p2[2]) - (p0[1] - p2[1]) * (p3[2] - p2[2]));
  dldx[1][1] = -t * ((p3[0] - p2[0]) * (p0[2] - p2[2]) - (p0[0] - p2[0]) * (p3[2] - p2[2]));
  dldx[1][2] = +t * ((p3[0] - p2[0]) * (p0[1] - p2[1]) - (p0[0] - p2[0]) * (p3[1] - p2[1]));
  dldx[2][0] = -t * ((p0[1] - p3[1]) * (p1[2] - p3[2]) - (p1[1] - p3[1]) * (p0[2] - p3[2]));
  dldx[2][1] = +t * ((p0[0] - p3[0]) * (p1[2] - p3[2]) - (p1[0] - p3[0]) * (p0[2] - p3[2]));
  dldx[2][2] = -t * ((p0[0] - p3[0]) * (p1[1] - p3[1]) - (p1[0] - p3[0]) * (p0[1] - p3[1]));
  dldx[3][0] = +t * ((p1[1] - p0[1]) * (p2[2] - p0[2]) - (p2[1] - p0[1]) * (p1[2] - p0[2]));
  dldx[3][1] = -t * ((p1[0] - p0[0]) * (p2[2] - p0[2]) - (p2[0] - p0[0]) * (p1[2] - p0[2]));
  dldx[3][2] = +t * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}
