// mlau_dev.h -- CDNA4 device functions of the MLAU per-step update.
//
// Point kernels shared by the 1D/2D/3D sweeps: limiters and WCNS face
// interpolation, the two characteristic reconstructions and the four Riemann
// solvers.  They are written for one wavefront lane each, in FP64, and keep
// the reference's evaluation order so that, compiled with
// -ffp-contract=off, every result equals the reference's (IEEE-correct
// division and sqrt, no FMA contraction).  The characteristic projections skip
// the eigenvector entries that are structurally zero or one; that changes at
// most the sign of an exact-zero intermediate, never a nonzero value.
//
// Reference anchors (file:line):
//   include/funcs.h:23-43        max/min (>= semantics), minmod, minmod3
//   include/interp.h:30-238      MUSCL (:30-37), WCNS3 (:106-146), WCNS4 (:186-238)
//   include/mhd_eigen.h:24-192   Athena eigensystem      (recon_char)
//   include/mhd_eigen.h:326-453  Alfven/entropy system   (recon_alfven)
//   include/mhd_eigen.h:455-684  mhd_c_reconst / mhd_a_reconst
//   include/mhd_flux.h:26-719    MLAU, Roe, HLLD, LHLLD
#pragma once
#include <hip/hip_runtime.h>

namespace mlau {

enum Flux { FLUX_MLAU = 0, FLUX_ROE = 1, FLUX_HLLD = 2, FLUX_LHLLD = 3 };

#define MLAU_DEV __device__ __forceinline__

MLAU_DEV double dmax(double a, double b) { return (a >= b) ? a : b; }
MLAU_DEV double dmin(double a, double b) { return (a >= b) ? b : a; }
MLAU_DEV double minmod(double a, double b) { return +dmin(dmax(a, b), 0.0) + dmax(dmin(a, b), 0.0); }

// include/interp.h:30-37
MLAU_DEV void muscl_mm(double f1, double f2, double f3, double &fl, double &fr)
{
  double df = 0.5 * minmod(f2 - f1, f3 - f2);
  fl = f2 + df;
  fr = f2 - df;
}

// include/interp.h:106-146
MLAU_DEV void wcns3(double f0, double f1, double f2, double &fl, double &fr)
{
  const double eps = 1e-40;
  double a = f0 - f1;
  double s0 = a * a;
  a = f2 - f1;
  double s1 = a * a;
  double t = f2 - 2.0 * f1 + f0;
  t = t * t;
  s0 = (1.0 + t / (s0 + eps));
  s1 = (1.0 + t / (s1 + eps));
  double w0 = 0.25 * s0, w1 = 0.75 * s1;
  w0 /= (w0 + w1);
  w1 = 1.0 - w0;
  fl = 0.5 * (w0 * (-f0 + 3.0 * f1) + w1 * (f1 + f2));
  w0 = 0.75 * s0;
  w1 = 0.25 * s1;
  w0 /= (w0 + w1);
  w1 = 1.0 - w0;
  fr = 0.5 * (w1 * (-f2 + 3.0 * f1) + w0 * (f1 + f0));
}

// include/interp.h:186-238
MLAU_DEV void wcns4(double f0, double f1, double f2, double f3, double f4, double &fl, double &fr)
{
  const double eps = 1e-40;
  double c1 = f0 - 4.0 * f1 + 3.0 * f2;
  double c2 = f0 - 2.0 * f1 + f2;
  double s0 = (13. / 12.) * c2 * c2 + 0.25 * c1 * c1;
  c1 = f1 - f3;
  c2 = f1 - 2.0 * f2 + f3;
  double s1 = (13. / 12.) * c2 * c2 + 0.25 * c1 * c1;
  c1 = f4 - 4.0 * f3 + 3.0 * f2;
  c2 = f4 - 2.0 * f3 + f2;
  double s2 = (13. / 12.) * c2 * c2 + 0.25 * c1 * c1;
  s0 = 1.0 / (s0 * s0 + eps);
  s1 = 1.0 / (s1 * s1 + eps);
  s2 = 1.0 / (s2 * s2 + eps);
  double w0 = 0.075 * s0, w1 = 0.650 * s1, w2 = 0.275 * s2;
  double d = 1.0 / (w0 + w1 + w2);
  w0 *= d;
  w1 *= d;
  w2 = 1.0 - w0 - w1;
  fl = 0.125 * (+w0 * (3.0 * f0 - 10.0 * f1 + 15.0 * f2) + w1 * (-f1 + 6.0 * f2 + 3.0 * f3) +
                w2 * (+3.0 * f2 + 6.0 * f3 - f4));
  w0 = 0.275 * s0;
  w1 = 0.650 * s1;
  w2 = 0.075 * s2;
  d = 1.0 / (w0 + w1 + w2);
  w0 *= d;
  w1 *= d;
  w2 = 1.0 - w0 - w1;
  fr = 0.125 * (+w2 * (3.0 * f4 - 10.0 * f3 + 15.0 * f2) + w1 * (-f3 + 6.0 * f2 + 3.0 * f1) +
                w0 * (+3.0 * f2 + 6.0 * f1 - f0));
}

// Face values of a 5-point stencil for spatial order ODR (solvers' calc_flr,
// 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:55-68).
template <int ODR>
MLAU_DEV void face_lr(const double (&f)[5], double &fl, double &fr)
{
  if constexpr (ODR == 1) {
    fl = f[2];
    fr = f[2];
  } else if constexpr (ODR == 2) {
    muscl_mm(f[1], f[2], f[3], fl, fr);
  } else if constexpr (ODR == 3) {
    wcns3(f[1], f[2], f[3], fl, fr);
  } else {
    wcns4(f[0], f[1], f[2], f[3], f[4], fl, fr);
  }
}

// Linear interpolation to faces (mhd_fd_ct_2d.c:42-57), used for the central
// part of the induction flux and for the corner E-field.
template <int ODR>
MLAU_DEV void lcal_lr(double f0, double f1, double f2, double f3, double f4, double &fl, double &fr)
{
  if constexpr (ODR == 1) {
    fl = f2;
    fr = f2;
  } else if constexpr (ODR == 2) {
    fl = 0.25 * (-f1 + 4.0 * f2 + f3);
    fr = 0.25 * (-f3 + 4.0 * f2 + f1);
  } else if constexpr (ODR == 3) {
    fl = 0.125 * (-f1 + 6.0 * f2 + 3.0 * f3);
    fr = 0.125 * (-f3 + 6.0 * f2 + 3.0 * f1);
  } else {
    fl = 0.003125 * (9.0 * f0 - 56.0 * f1 + 234.0 * f2 + 144.0 * f3 - 11.0 * f4);
    fr = 0.003125 * (9.0 * f4 - 56.0 * f3 + 234.0 * f2 + 144.0 * f1 - 11.0 * f0);
  }
}
// Left value only (at +1/2 of the centre) and right value only (at -1/2).
template <int ODR>
MLAU_DEV double lcal_l(double f0, double f1, double f2, double f3, double f4)
{
  if constexpr (ODR == 1) return f2;
  else if constexpr (ODR == 2) return 0.25 * (-f1 + 4.0 * f2 + f3);
  else if constexpr (ODR == 3) return 0.125 * (-f1 + 6.0 * f2 + 3.0 * f3);
  else return 0.003125 * (9.0 * f0 - 56.0 * f1 + 234.0 * f2 + 144.0 * f3 - 11.0 * f4);
}
template <int ODR>
MLAU_DEV double lcal_r(double f0, double f1, double f2, double f3, double f4)
{
  if constexpr (ODR == 1) return f2;
  else if constexpr (ODR == 2) return 0.25 * (-f3 + 4.0 * f2 + f1);
  else if constexpr (ODR == 3) return 0.125 * (-f3 + 6.0 * f2 + 3.0 * f1);
  else return 0.003125 * (9.0 * f4 - 56.0 * f3 + 234.0 * f2 + 144.0 * f1 - 11.0 * f0);
}

// Face B -> cell centre (mhd_fd_ct_2d.c:32-40)
template <int ODR>
MLAU_DEV double bc_centre(double f0, double f1, double f2, double f3)
{
  if constexpr (ODR <= 2) return 0.5 * (f1 + f2);
  else return 0.0625 * (9.0 * (f1 + f2) - (f0 + f3));
}
// First difference (mhd_fd_ct_2d.c:74-82)
template <int ODR>
MLAU_DEV double diff4(double f0, double f1, double f2, double f3)
{
  if constexpr (ODR <= 2) return f2 - f1;
  else return (27.0 * (f2 - f1) - (f3 - f0)) / 24.0;
}

// Stencil of the seven rotated primitives (ro,vn,vt,vu,bt,bu,pr) at 5 points.
struct Stencil {
  double q[7][5];
};

// Exact-zero forcing of include/mhd_eigen.h:516-529
MLAU_DEV void force_zeros(const Stencil &S, double (&ul)[7], double (&ur)[7])
{
#pragma unroll
  for (int v = 1; v <= 5; v++) {
    bool z = (S.q[v][0] == 0.0) & (S.q[v][1] == 0.0) & (S.q[v][2] == 0.0) & (S.q[v][3] == 0.0) &
             (S.q[v][4] == 0.0);
    if (z) {
      ul[v] = 0.0;
      ur[v] = 0.0;
    }
  }
}

// Characteristic reconstruction with the Athena eigensystem:
// mhd_c_reconst + mhd_eigen_st (include/mhd_eigen.h:455-530, 24-192).
template <int ODR>
MLAU_DEV void recon_char(const Stencil &S, double bx, double gamma, double (&ul)[7], double (&ur)[7])
{
  const double ro = S.q[0][2], by = S.q[4][2], bz = S.q[5][2], pr = S.q[6][2];
  const double eps = 1e-12;
  const double hsqrt2 = 0.5 * sqrt(2.0);
  double iro = 1.0 / ro;
  double sro = sqrt(ro);
  double isro = 1.0 / sro;
  double bxsr = bx * isro, bysr = by * isro, bzsr = bz * isro;
  double bx2 = bxsr * bxsr;
  double bt2 = bysr * bysr + bzsr * bzsr;
  double b2 = bx2 + bt2;
  double a2 = fabs(gamma * pr / ro);
  double aa = sqrt(a2);
  double ca2 = bx2;
  double junk = sqrt(fabs((a2 + b2) * (a2 + b2) - 4.0 * a2 * ca2));
  double cf2 = fabs(0.5 * (a2 + b2 + junk));
  double cf = sqrt(cf2);
  double cs2 = fabs(0.5 * (a2 + b2 - junk));
  double cs = sqrt(cs2);
  bx2 = bx * bx;
  bt2 = by * by + bz * bz;
  b2 = bx2 + bt2;
  double den = 1.0 / sqrt(bt2);
  bool flag = (bt2 > eps * b2);
  double betay = flag ? (by * den) : hsqrt2;
  double betaz = flag ? (bz * den) : hsqrt2;
  den = 1.0 / sqrt(fabs(cf2 - cs2));
  flag = ((bt2 > eps * b2) || (fabs(gamma * pr - bx2) > eps * gamma * pr));
  double alphaf = flag ? (sqrt(fabs(a2 - cs2)) * den) : 1.0;
  double alphas = flag ? (sqrt(fabs(cf2 - a2)) * den) : 0.0;
  double sgnbx = (bx > 0) ? (1.0) : (-1.0);
  double cfalf = cf * alphaf, csals = cs * alphas, aaalf = aa * alphaf, aaals = aa * alphas;
  den = 0.5 / a2;
  // left eigenvectors l1, l2, l3 (l5..l7 are mirrors, l4 = (1,0,0,0,0,0,-1/a2))
  const double l11 = -cfalf * den, l12 = +csals * den * sgnbx * betay, l13 = +csals * den * sgnbx * betaz;
  const double l14 = +aaals * den * isro * betay, l15 = +aaals * den * isro * betaz, l16 = alphaf * den * iro;
  const double l22 = -0.5 * betaz, l23 = +0.5 * betay, l24 = -0.5 * isro * betaz * sgnbx;
  const double l25 = +0.5 * isro * betay * sgnbx;
  const double l31 = -csals * den, l32 = -cfalf * den * sgnbx * betay, l33 = -cfalf * den * sgnbx * betaz;
  const double l34 = -aaalf * den * isro * betay, l35 = -aaalf * den * isro * betaz, l36 = alphas * den * iro;
  const double l46 = -1.0 / a2;
  double w[7][5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const double r_ = S.q[0][m], x_ = S.q[1][m], y_ = S.q[2][m], z_ = S.q[3][m];
    const double b_ = S.q[4][m], c_ = S.q[5][m], p_ = S.q[6][m];
    w[0][m] = l11 * x_ + l12 * y_ + l13 * z_ + l14 * b_ + l15 * c_ + l16 * p_;
    w[1][m] = l22 * y_ + l23 * z_ + l24 * b_ + l25 * c_;
    w[2][m] = l31 * x_ + l32 * y_ + l33 * z_ + l34 * b_ + l35 * c_ + l36 * p_;
    w[3][m] = r_ + l46 * p_;
    w[4][m] = (-l31) * x_ + (-l32) * y_ + (-l33) * z_ + l34 * b_ + l35 * c_ + l36 * p_;
    w[5][m] = (-l22) * y_ + (-l23) * z_ + l24 * b_ + l25 * c_;
    w[6][m] = (-l11) * x_ + (-l12) * y_ + (-l13) * z_ + l14 * b_ + l15 * c_ + l16 * p_;
  }
  double wl[7], wr[7];
#pragma unroll
  for (int n = 0; n < 7; n++) face_lr<ODR>(w[n], wl[n], wr[n]);
  // right eigenvectors r1, r2, r3 (r5..r7 mirrors, r4 = e_0)
  const double r10 = ro * alphaf, r11 = -cfalf, r12 = +csals * sgnbx * betay, r13 = +csals * sgnbx * betaz;
  const double r14 = +aaals * sro * betay, r15 = +aaals * sro * betaz, r16 = alphaf * gamma * pr;
  const double r22 = -betaz, r23 = +betay, r24 = -sro * betaz * sgnbx, r25 = +sro * betay * sgnbx;
  const double r30 = ro * alphas, r31 = -csals, r32 = -cfalf * sgnbx * betay, r33 = -cfalf * sgnbx * betaz;
  const double r34 = -aaalf * sro * betay, r35 = -aaalf * sro * betaz, r36 = alphas * gamma * pr;
#define MLAU_BACKPROJ(W, U)                                                                         \
  U[0] = r10 * W[0] + r30 * W[2] + W[3] + r30 * W[4] + r10 * W[6];                                  \
  U[1] = r11 * W[0] + r31 * W[2] + (-r31) * W[4] + (-r11) * W[6];                                   \
  U[2] = r12 * W[0] + r22 * W[1] + r32 * W[2] + (-r32) * W[4] + (-r22) * W[5] + (-r12) * W[6];      \
  U[3] = r13 * W[0] + r23 * W[1] + r33 * W[2] + (-r33) * W[4] + (-r23) * W[5] + (-r13) * W[6];      \
  U[4] = r14 * W[0] + r24 * W[1] + r34 * W[2] + r34 * W[4] + r24 * W[5] + r14 * W[6];               \
  U[5] = r15 * W[0] + r25 * W[1] + r35 * W[2] + r35 * W[4] + r25 * W[5] + r15 * W[6];               \
  U[6] = r16 * W[0] + r36 * W[2] + r36 * W[4] + r16 * W[6];
  MLAU_BACKPROJ(wl, ul)
  MLAU_BACKPROJ(wr, ur)
#undef MLAU_BACKPROJ
  force_zeros(S, ul, ur);
}

// Alfven/entropy reconstruction: mhd_a_reconst + mhd_eigen_ae
// (include/mhd_eigen.h:609-684, 326-453).
template <int ODR>
MLAU_DEV void recon_alfven(const Stencil &S, double bx, double gamma, double (&ul)[7], double (&ur)[7])
{
  const double ro = S.q[0][2], by = S.q[4][2], bz = S.q[5][2], pr = S.q[6][2];
  double iro = 1.0 / ro;
  double cs2 = gamma * pr * iro;
  double ca2 = bx * bx * iro;
  double ct2 = (by * by + bz * bz) * iro;
  double cb2 = cs2 + ca2 + ct2;
  double cf2 = 0.5 * (cb2 + sqrt(fabs(cb2 * cb2 - 4.0 * cs2 * ca2)));
  double sro = sqrt(ro);
  double isroh = 0.5 / sro;
  double icf2 = 1.0 / cf2;
  double byh = 0.5 * by, bzh = 0.5 * bz;
  double w[7][5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const double r_ = S.q[0][m], x_ = S.q[1][m], y_ = S.q[2][m], z_ = S.q[3][m];
    const double b_ = S.q[4][m], c_ = S.q[5][m], p_ = S.q[6][m];
    w[0][m] = x_;
    w[1][m] = cf2 * r_ + (-by) * b_ + (-bz) * c_ + (-p_);
    w[2][m] = by * b_ + bz * c_ + p_;
    w[3][m] = sro * y_ + b_;
    w[4][m] = sro * y_ + (-b_);
    w[5][m] = sro * z_ + c_;
    w[6][m] = sro * z_ + (-c_);
  }
  double wl[7], wr[7];
#pragma unroll
  for (int n = 0; n < 7; n++) face_lr<ODR>(w[n], wl[n], wr[n]);
#define MLAU_BACKPROJ(W, U)                                                  \
  U[0] = icf2 * W[1] + icf2 * W[2];                                          \
  U[1] = W[0];                                                               \
  U[2] = isroh * W[3] + isroh * W[4];                                        \
  U[3] = isroh * W[5] + isroh * W[6];                                        \
  U[4] = 0.5 * W[3] + (-0.5) * W[4];                                         \
  U[5] = 0.5 * W[5] + (-0.5) * W[6];                                         \
  U[6] = W[2] + (-byh) * W[3] + byh * W[4] + (-bzh) * W[5] + bzh * W[6];
  MLAU_BACKPROJ(wl, ul)
  MLAU_BACKPROJ(wr, ur)
#undef MLAU_BACKPROJ
  force_zeros(S, ul, ur);
}

#include "mlau_riemann.inc"

// Reconstruction of one cell: characteristic (Athena) projection when the
// stencil is supersonic, Alfven/entropy projection otherwise.  The two paths
// only differ in their eigenvectors, so the seven face interpolations (the
// bulk of the work) are shared: a wavefront holding both kinds of cells runs
// the projections under divergence but the WCNS/MUSCL part once.
// Bit-identical to recon_char / recon_alfven above.
template <int ODR>
MLAU_DEV void recon(const Stencil &S, bool supersonic, double bx, double gamma, double (&ul)[7], double (&ur)[7])
{
  const double ro = S.q[0][2], by = S.q[4][2], bz = S.q[5][2], pr = S.q[6][2];
  double w[7][5];
  // back-projection coefficients (meaning depends on the path)
  double r10, r11, r12, r13, r14, r15, r16, r22, r23, r24, r25, r30, r31, r32, r33, r34, r35, r36;
  if (supersonic) {
    const double eps = 1e-12;
    const double hsqrt2 = 0.5 * sqrt(2.0);
    double iro = 1.0 / ro;
    double sro = sqrt(ro);
    double isro = 1.0 / sro;
    double bxsr = bx * isro, bysr = by * isro, bzsr = bz * isro;
    double bx2 = bxsr * bxsr;
    double bt2 = bysr * bysr + bzsr * bzsr;
    double b2 = bx2 + bt2;
    double a2 = fabs(gamma * pr / ro);
    double aa = sqrt(a2);
    double ca2 = bx2;
    double junk = sqrt(fabs((a2 + b2) * (a2 + b2) - 4.0 * a2 * ca2));
    double cf2 = fabs(0.5 * (a2 + b2 + junk));
    double cf = sqrt(cf2);
    double cs2 = fabs(0.5 * (a2 + b2 - junk));
    double cs = sqrt(cs2);
    bx2 = bx * bx;
    bt2 = by * by + bz * bz;
    b2 = bx2 + bt2;
    double den = 1.0 / sqrt(bt2);
    bool flag = (bt2 > eps * b2);
    double betay = flag ? (by * den) : hsqrt2;
    double betaz = flag ? (bz * den) : hsqrt2;
    den = 1.0 / sqrt(fabs(cf2 - cs2));
    flag = ((bt2 > eps * b2) || (fabs(gamma * pr - bx2) > eps * gamma * pr));
    double alphaf = flag ? (sqrt(fabs(a2 - cs2)) * den) : 1.0;
    double alphas = flag ? (sqrt(fabs(cf2 - a2)) * den) : 0.0;
    double sgnbx = (bx > 0) ? (1.0) : (-1.0);
    double cfalf = cf * alphaf, csals = cs * alphas, aaalf = aa * alphaf, aaals = aa * alphas;
    den = 0.5 / a2;
    const double l11 = -cfalf * den, l12 = +csals * den * sgnbx * betay, l13 = +csals * den * sgnbx * betaz;
    const double l14 = +aaals * den * isro * betay, l15 = +aaals * den * isro * betaz, l16 = alphaf * den * iro;
    const double l22 = -0.5 * betaz, l23 = +0.5 * betay, l24 = -0.5 * isro * betaz * sgnbx;
    const double l25 = +0.5 * isro * betay * sgnbx;
    const double l31 = -csals * den, l32 = -cfalf * den * sgnbx * betay, l33 = -cfalf * den * sgnbx * betaz;
    const double l34 = -aaalf * den * isro * betay, l35 = -aaalf * den * isro * betaz, l36 = alphas * den * iro;
    const double l46 = -1.0 / a2;
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const double r_ = S.q[0][m], x_ = S.q[1][m], y_ = S.q[2][m], z_ = S.q[3][m];
      const double b_ = S.q[4][m], c_ = S.q[5][m], p_ = S.q[6][m];
      w[0][m] = l11 * x_ + l12 * y_ + l13 * z_ + l14 * b_ + l15 * c_ + l16 * p_;
      w[1][m] = l22 * y_ + l23 * z_ + l24 * b_ + l25 * c_;
      w[2][m] = l31 * x_ + l32 * y_ + l33 * z_ + l34 * b_ + l35 * c_ + l36 * p_;
      w[3][m] = r_ + l46 * p_;
      w[4][m] = (-l31) * x_ + (-l32) * y_ + (-l33) * z_ + l34 * b_ + l35 * c_ + l36 * p_;
      w[5][m] = (-l22) * y_ + (-l23) * z_ + l24 * b_ + l25 * c_;
      w[6][m] = (-l11) * x_ + (-l12) * y_ + (-l13) * z_ + l14 * b_ + l15 * c_ + l16 * p_;
    }
    r10 = ro * alphaf; r11 = -cfalf; r12 = +csals * sgnbx * betay; r13 = +csals * sgnbx * betaz;
    r14 = +aaals * sro * betay; r15 = +aaals * sro * betaz; r16 = alphaf * gamma * pr;
    r22 = -betaz; r23 = +betay; r24 = -sro * betaz * sgnbx; r25 = +sro * betay * sgnbx;
    r30 = ro * alphas; r31 = -csals; r32 = -cfalf * sgnbx * betay; r33 = -cfalf * sgnbx * betaz;
    r34 = -aaalf * sro * betay; r35 = -aaalf * sro * betaz; r36 = alphas * gamma * pr;
  } else {
    double iro = 1.0 / ro;
    double cs2 = gamma * pr * iro;
    double ca2 = bx * bx * iro;
    double ct2 = (by * by + bz * bz) * iro;
    double cb2 = cs2 + ca2 + ct2;
    double cf2 = 0.5 * (cb2 + sqrt(fabs(cb2 * cb2 - 4.0 * cs2 * ca2)));
    double sro = sqrt(ro);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const double r_ = S.q[0][m], x_ = S.q[1][m], y_ = S.q[2][m], z_ = S.q[3][m];
      const double b_ = S.q[4][m], c_ = S.q[5][m], p_ = S.q[6][m];
      w[0][m] = x_;
      w[1][m] = cf2 * r_ + (-by) * b_ + (-bz) * c_ + (-p_);
      w[2][m] = by * b_ + bz * c_ + p_;
      w[3][m] = sro * y_ + b_;
      w[4][m] = sro * y_ + (-b_);
      w[5][m] = sro * z_ + c_;
      w[6][m] = sro * z_ + (-c_);
    }
    r10 = 1.0 / cf2;        // icf2
    r11 = 0.5 / sro;        // isroh
    r12 = 0.5 * by;         // byh
    r13 = 0.5 * bz;         // bzh
    r14 = r15 = r16 = r22 = r23 = r24 = r25 = r30 = r31 = r32 = r33 = r34 = r35 = r36 = 0.0;
  }
  double wl[7], wr[7];
#pragma unroll
  for (int n = 0; n < 7; n++) face_lr<ODR>(w[n], wl[n], wr[n]);
  if (supersonic) {
#define MLAU_BACKPROJ(W, U)                                                                         \
  U[0] = r10 * W[0] + r30 * W[2] + W[3] + r30 * W[4] + r10 * W[6];                                  \
  U[1] = r11 * W[0] + r31 * W[2] + (-r31) * W[4] + (-r11) * W[6];                                   \
  U[2] = r12 * W[0] + r22 * W[1] + r32 * W[2] + (-r32) * W[4] + (-r22) * W[5] + (-r12) * W[6];      \
  U[3] = r13 * W[0] + r23 * W[1] + r33 * W[2] + (-r33) * W[4] + (-r23) * W[5] + (-r13) * W[6];      \
  U[4] = r14 * W[0] + r24 * W[1] + r34 * W[2] + r34 * W[4] + r24 * W[5] + r14 * W[6];               \
  U[5] = r15 * W[0] + r25 * W[1] + r35 * W[2] + r35 * W[4] + r25 * W[5] + r15 * W[6];               \
  U[6] = r16 * W[0] + r36 * W[2] + r36 * W[4] + r16 * W[6];
    MLAU_BACKPROJ(wl, ul)
    MLAU_BACKPROJ(wr, ur)
#undef MLAU_BACKPROJ
  } else {
    const double icf2 = r10, isroh = r11, byh = r12, bzh = r13;
#define MLAU_BACKPROJ(W, U)                                                  \
  U[0] = icf2 * W[1] + icf2 * W[2];                                          \
  U[1] = W[0];                                                               \
  U[2] = isroh * W[3] + isroh * W[4];                                        \
  U[3] = isroh * W[5] + isroh * W[6];                                        \
  U[4] = 0.5 * W[3] + (-0.5) * W[4];                                         \
  U[5] = 0.5 * W[5] + (-0.5) * W[6];                                         \
  U[6] = W[2] + (-byh) * W[3] + byh * W[4] + (-bzh) * W[5] + bzh * W[6];
    MLAU_BACKPROJ(wl, ul)
    MLAU_BACKPROJ(wr, ur)
#undef MLAU_BACKPROJ
  }
  force_zeros(S, ul, ur);
}

}  // namespace mlau
