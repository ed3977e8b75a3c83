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
//   include/interp.h:76-284      MUSCL, WCNS3, WCNS4
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

// include/interp.h:76-83
MLAU_DEV void muscl_mm(double f1, double f2, double f3, double &fl, double &fr)
{
  double df = 0.5 * minmod(f2 - f1, f3 - f2);
  fl = f2 + df;
  fr = f2 - df;
}

// include/interp.h:152-192
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

// include/interp.h:232-284
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

// ------------------------------------------------------------ Riemann solvers
// Inputs: left (ro,vn,vt,vu,bt,bu,pr), right likewise, normal B, gamma,
// shock-detector dv[2].  Output (ro,mn,mt,mu,bt,bu,en) fluxes.

// include/mhd_flux.h:26-213
MLAU_DEV void flux_mlau(const double (&L)[7], const double (&R)[7], double bnc, double gamma, double dv0,
                        double dv1, double (&f)[7])
{
  const double rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const double ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  const double EPS = 1e-8;
  double gammam1i = 1.0 / (gamma - 1.0);
  double bnc2 = bnc * bnc;
  double sgn = (bnc > 0) ? (1.0) : (-1.0);
  double bnca = sgn * bnc;
  double roli = 1.0 / rol;
  double vl2 = vnl * vnl + vtl * vtl + vul * vul;
  double pml = 0.5 * (btl * btl + bul * bul);
  double ptl = prl + pml;
  double enl = gammam1i * prl + pml + 0.5 * rol * vl2;
  double hnl = (enl + ptl) * roli;
  double rori = 1.0 / ror;
  double vr2 = vnr * vnr + vtr * vtr + vur * vur;
  double pmr = 0.5 * (btr * btr + bur * bur);
  double ptr = prr + pmr;
  double enr = gammam1i * prr + pmr + 0.5 * ror * vr2;
  double hnr = (enr + ptr) * rori;
  double cl2 = gamma * prl * roli;
  double cr2 = gamma * prr * rori;
  double cal2 = bnc2 * roli;
  double car2 = bnc2 * rori;
  double cbl2 = cl2 + cal2 + 2.0 * pml * roli;
  double cbr2 = cr2 + car2 + 2.0 * pmr * rori;
  double cfl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * cl2 * cal2)));
  double cfr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * cr2 * car2)));
  cbl2 = vl2 + cal2 + 2.0 * pml * roli;
  cbr2 = vr2 + car2 + 2.0 * pmr * rori;
  double ccl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * vl2 * cal2)));
  double ccr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * vr2 * car2)));
  double cmax = sqrt(dmax(cfl2, cfr2));
  double cfn = cmax;
  double cfni = 1.0 / cfn;
  double sl = dmin(0.0, dmin(vnl, vnr) - cmax);
  double sr = dmax(0.0, dmax(vnl, vnr) + cmax);
  double mal = vnl * cfni;
  double mar = vnr * cfni;
  double snl = (mal > 0) ? 1.0 : (-1.0);
  double snr = (mar > 0) ? 1.0 : (-1.0);
  double theta = dmin(1.0, (cfn - dmin(dv0, 0.0)) / (cfn - dmin(dv1, 0.0)));
  theta *= theta;
  theta *= theta;
  const double alpha = 0.1875, beta = 0.125;
  double malp1 = mal + 1.0;
  double marm1 = mar - 1.0;
  double mal2m1 = mal * mal - 1.0;
  double mar2m1 = mar * mar - 1.0;
  double bfl, ffl, bfr, ffr;
  if (snl * mal > 1.0) {
    bfl = 0.5 * (1.0 + snl);
    ffl = bfl * mal;
  } else {
    bfl = +0.25 * malp1 * malp1 * (2.0 - mal) + alpha * mal * mal2m1 * mal2m1;
    ffl = +0.25 * malp1 * malp1 + beta * mal2m1 * mal2m1;
  }
  if (snr * mar > 1.0) {
    bfr = 0.5 * (1.0 - snr);
    ffr = bfr * mar;
  } else {
    bfr = +0.25 * marm1 * marm1 * (2.0 + mar) - alpha * mar * mar2m1 * mar2m1;
    ffr = -0.25 * marm1 * marm1 - beta * mar2m1 * mar2m1;
  }
  double mnc = ffl + ffr;
  double mflux = mnc * cfn;
  mflux -= dmax(1.0 - fabs(mnc), 0.0) * theta * cfni * (ptr - ptl) / (rol + ror);
  double dir = (mflux > 0) ? 1.0 : (-1.0);
  double dirl = 0.5 * (1.0 + dir);
  double dirr = 1.0 - dirl;
  mflux *= (dirl * rol + dirr * ror);
  double mfl = mflux * dirl;
  double mfr = mflux - mfl;
  double ptot = bfl * ptl + bfr * ptr;
  double cfmod = sqrt(dmax(ccl2, ccr2));
  ptot -= 0.25 * bfl * bfr * (rol + ror) * cfmod * (vnr - vnl);
  ptot -= 0.5 * (1.0 - cfmod * cfni) * (bfl + bfr - 1.0) * (ptl + ptr);
  double slvl = sl - vnl;
  double srvr = sr - vnr;
  double snc = dirl * sl + dirr * sr;
  double vndnm = mflux + dirl * rol * slvl + dirr * ror * srvr;
  double vnc = ((fabs(vndnm) > EPS) && (snc != 0)) ? (mflux * snc / vndnm) : (dirl * vnl + dirr * vnr);
  double slvc = sl - vnc;
  double srvc = sr - vnc;
  double slvci = 1.0 / slvc;
  double srvci = 1.0 / srvc;
  double rhdl = rol * slvl * slvc - bnc2;
  double rhdr = ror * srvr * srvc - bnc2;
  double dvtl = 0.0, dvul = 0.0, dbtl = 0.0, dbul = 0.0;
  double dvtr = 0.0, dvur = 0.0, dbtr = 0.0, dbur = 0.0;
  if (fabs(rhdl) > EPS) {
    double q = (vnc - vnl) / rhdl;
    dvtl = -bnc * btl * q;
    dvul = -bnc * bul * q;
    dbtl = bnc2 * btl * q * slvci;
    dbul = bnc2 * bul * q * slvci;
  }
  if (fabs(rhdr) > EPS) {
    double q = (vnc - vnr) / rhdr;
    dvtr = -bnc * btr * q;
    dvur = -bnc * bur * q;
    dbtr = bnc2 * btr * q * srvci;
    dbur = bnc2 * bur * q * srvci;
  }
  double cmpl = slvl * slvci;
  double cmpr = srvr * srvci;
  double ro2l = rol * cmpl;
  double srol = sqrt(ro2l);
  double vt2l = dvtl + vtl;
  double vu2l = dvul + vul;
  double bt2l = dbtl + btl * cmpl;
  double bu2l = dbul + bul * cmpl;
  double ro2r = ror * cmpr;
  double sror = sqrt(ro2r);
  double vt2r = dvtr + vtr;
  double vu2r = dvur + vur;
  double bt2r = dbtr + btr * cmpr;
  double bu2r = dbur + bur * cmpr;
  double sroi = 1.0 / (srol + sror);
  double srollri = srol * sroi;
  double srorlri = sror * sroi;
  double bnflb = sgn * dmin(dmax(0.0, srollri * (bnca + sror * vnc)), bnca);
  double bnfrb = sgn * dmin(dmax(0.0, srorlri * (bnca - srol * vnc)), bnca);
  double bnflv = sgn * dmin(dmax(0.0, srorlri * (bnca + mflux / sror)), bnca);
  double bnfrv = sgn * dmin(dmax(0.0, srollri * (bnca - mflux / srol)), bnca);
  double bnsrv = dmax(bnca - dir * vnc * (dirl * srol + dirr * sror), 0.0);
  double ccb = sroi * bnsrv;
  double ccv = srol * sror * ccb;
  double bnvt = bnflb * vt2l + bnfrb * vt2r + ccb * (bt2r - bt2l);
  double bnvu = bnflb * vu2l + bnfrb * vu2r + ccb * (bu2r - bu2l);
  double bnbt = bnflv * bt2l + bnfrv * bt2r + ccv * (vt2r - vt2l);
  double bnbu = bnflv * bu2l + bnfrv * bu2r + ccv * (vu2r - vu2l);
  bnvt -= vnc * (dirl * dbtl + dirr * dbtr);
  bnvu -= vnc * (dirl * dbul + dirr * dbur);
  bnbt -= mflux * (dirl * dvtl + dirr * dvtr);
  bnbu -= mflux * (dirl * dvul + dirr * dvur);
  double fmt = mfl * vtl + mfr * vtr - bnbt;
  double fmu = mfl * vul + mfr * vur - bnbu;
  double fbt = mfl * btl * roli + mfr * btr * rori - bnvt;
  double fbu = mfl * bul * roli + mfr * bur * rori - bnvu;
  f[0] = mflux;
  f[1] = mfl * vnl + mfr * vnr + ptot - 0.5 * bnc2;
  f[2] = fmt;
  f[3] = fmu;
  f[4] = fbt;
  f[5] = fbu;
  double vb1 = dirl * (vtl * btl + vul * bul) + dirr * (vtr * btr + vur * bur);
  double vb2 = dirl * (vt2l * bt2l + vu2l * bu2l) + dirr * (vt2r * bt2r + vu2r * bu2r);
  double rv2b2 = (mflux * vnc - bnc2);
  rv2b2 *= rv2b2;
  double vb3 = 0.0;
  if (rv2b2 > EPS * EPS) {
    vb3 = (+(fmt * vnc + fbt * bnc) * (fmt * bnc + fbt * mflux) + (fmu * vnc + fbu * bnc) * (fmu * bnc + fbu * mflux)) /
          rv2b2;
  }
  double bnvb = sgn * (bnca * (snc * vb2 - vnc * vb1) * (dirl * slvci + dirr * srvci) + bnsrv * (vb3 - vb2));
  f[6] = mfl * hnl + mfr * hnr - bnvb;
}

// include/mhd_flux.h:215-454
MLAU_DEV void flux_roe(const double (&L)[7], const double (&R)[7], double bnc, double gamma, double (&f)[7])
{
  const double rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const double ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  const double EPS = 1e-8;
  double gammam1 = gamma - 1.0;
  double gammam2 = gamma - 2.0;
  double gammam1i = 1.0 / gammam1;
  double roli = 1.0 / rol;
  double rori = 1.0 / ror;
  double rrol = sqrt(rol);
  double rror = sqrt(ror);
  double rroi = 1.0 / (rrol + rror);
  double robar = rrol * rror;
  double robari = 1.0 / robar;
  double rrobar = sqrt(robar);
  double rrobai = sqrt(robari);
  double vnbar = (rrol * vnl + rror * vnr) * rroi;
  double vtbar = (rrol * vtl + rror * vtr) * rroi;
  double vubar = (rrol * vul + rror * vur) * rroi;
  double bnbar = bnc;
  double btbar = (rrol * btr + rror * btl) * rroi;
  double bubar = (rrol * bur + rror * bul) * rroi;
  double hl = (+gamma * gammam1i * prl * roli + 0.5 * (vnl * vnl + vtl * vtl + vul * vul) +
               (bnbar * bnbar + btl * btl + bul * bul) * roli);
  double hr = (+gamma * gammam1i * prr * rori + 0.5 * (vnr * vnr + vtr * vtr + vur * vur) +
               (bnbar * bnbar + btr * btr + bur * bur) * rori);
  double hbar = (rrol * hl + rror * hr) * rroi;
  double btave = 0.5 * (btl + btr);
  double buave = 0.5 * (bul + bur);
  double sgn = (bnbar > 0) ? (1.0) : (-1.0);
  double db2 = 0.5 * gammam2 * gammam1i * ((btr - btl) * (btr - btl) + (bur - bul) * (bur - bul)) * rroi * rroi;
  double v2h = 0.5 * (vnbar * vnbar + vtbar * vtbar + vubar * vubar);
  double bb = (bnbar * bnbar + btbar * btbar + bubar * bubar);
  double c2 = gammam1 * (+hbar - v2h - db2 - bb * robari);
  double a2 = (gammam1 * (+hbar - v2h - db2) - gammam2 * bb * robari);
  double ca2 = bnbar * bnbar * robari;
  double cb2 = (btbar * btbar + bubar * bubar) * robari;
  double cf2 = 0.5 * (a2 + sqrt(cb2 * (a2 + c2 + ca2) + (c2 - ca2) * (c2 - ca2)));
  double cs2 = c2 * ca2 / cf2;
  double cf = sqrt(cf2);
  double cs = sqrt(cs2);
  double ca = sqrt(ca2);
  double c = sqrt(c2);
  double cfi = 1.0 / cf;
  double sgr1 = btbar * btbar + bubar * bubar - EPS;
  double sf11 = (sgr1 > 0) ? 1.0 : 0.0;
  double sf21 = 1.0 - sf11;
  double beti = 1.0 / sqrt(btbar * btbar + bubar * bubar + sf21);
  const double sqh = sqrt(0.5);
  double bett = sf11 * btbar * beti + sf21 * sqh;
  double betu = sf11 * bubar * beti + sf21 * sqh;
  double sgr2 = (btbar * btbar + bubar * bubar) * robari + fabs(ca2 - c2) - EPS;
  double sf12 = (sgr2 > 0) ? 1.0 : 0.0;
  double sf22 = 1.0 - sf12;
  double cfca = dmax(0.0, cf2 - ca2);
  double cfcs = dmax(0.0, cf2 - cs2);
  double cfa = dmax(0.0, cf2 - c2);
  double alpi = 1.0 / sqrt(cfcs + sf22);
  double alpf = sf12 * sqrt(cfca) * alpi + sf22;
  double alps = sf12 * sqrt(cfa) * alpi;
  const double eeps = 0.0;
  double evpf = dmax(fabs(vnbar + cf), eeps);
  double evmf = dmax(fabs(vnbar - cf), eeps);
  double evps = dmax(fabs(vnbar + cs), eeps);
  double evms = dmax(fabs(vnbar - cs), eeps);
  double evpa = dmax(fabs(vnbar + ca), eeps);
  double evma = dmax(fabs(vnbar - ca), eeps);
  double ev00 = dmax(fabs(vnbar), eeps);
  double dltro = ror - rol;
  double dltmn = robar * (vnr - vnl);
  double dltmt = robar * (vtr - vtl);
  double dltmu = robar * (vur - vul);
  double dltbt = btr - btl;
  double dltbu = bur - bul;
  double t1 = bett * dltbt + betu * dltbu;
  double t2 = (prr - prl + (btave * dltbt + buave * dltbu) + gammam2 * (btbar * dltbt + bubar * dltbu)) * gammam1i;
  double t3 = betu * dltbt - bett * dltbu;
  double s1 = dltmn;
  double s2 = bett * dltmt + betu * dltmu;
  double s3 = betu * dltmt - bett * dltmu;
  double p11 = alps * cf * rrobai;
  double p12 = -alpf * c2 * cfi * rrobai;
  double p21 = alpf * (cf2 - c2 * gammam2 * gammam1i);
  double p22 = alps * (cs2 - c2 * gammam2 * gammam1i);
  double q11 = alpf * cf;
  double q12 = alps * cs;
  double q21 = -alps * ca * sgn;
  double q22 = alpf * c * sgn;
  double detpi = 1.0 / (p11 * p22 - p12 * p21);
  double detqi = 1.0 / (q11 * q22 - q12 * q21);
  double wvpf = 0.5 * ((p22 * t1 - p12 * t2) * detpi + (q22 * s1 - q12 * s2) * detqi);
  double wvmf = 0.5 * ((p22 * t1 - p12 * t2) * detpi - (q22 * s1 - q12 * s2) * detqi);
  double wvps = 0.5 * ((-p21 * t1 + p11 * t2) * detpi + (-q21 * s1 + q11 * s2) * detqi);
  double wvms = 0.5 * ((-p21 * t1 + p11 * t2) * detpi - (-q21 * s1 + q11 * s2) * detqi);
  double wvpa = 0.5 * (rrobar * t3 - sgn * s3);
  double wvma = 0.5 * (rrobar * t3 + sgn * s3);
  double wv00 = dltro - alpf * (wvpf + wvmf) - alps * (wvps + wvms);
  double frol = rol * vnl;
  double fmnl = rol * vnl * vnl + prl + 0.5 * (-bnbar * bnbar + btl * btl + bul * bul);
  double fmtl = rol * vnl * vtl - bnbar * btl;
  double fmul = rol * vnl * vul - bnbar * bul;
  double fbtl = vnl * btl - vtl * bnbar;
  double fbul = vnl * bul - vul * bnbar;
  double fenl = rol * vnl * hl - bnbar * (bnbar * vnl + btl * vtl + bul * vul);
  double fror = ror * vnr;
  double fmnr = ror * vnr * vnr + prr + 0.5 * (-bnbar * bnbar + btr * btr + bur * bur);
  double fmtr = ror * vnr * vtr - bnbar * btr;
  double fmur = ror * vnr * vur - bnbar * bur;
  double fbtr = vnr * btr - vtr * bnbar;
  double fbur = vnr * bur - vur * bnbar;
  double fenr = ror * vnr * hr - bnbar * (bnbar * vnr + btr * vtr + bur * vur);
  double vb = (bett * vtbar + betu * vubar);
  double rvpfro = alpf;
  double rvpfmn = alpf * (vnbar + cf);
  double rvpfmt = alpf * vtbar - alps * bett * ca * sgn;
  double rvpfmu = alpf * vubar - alps * betu * ca * sgn;
  double rvpfbt = alps * bett * cf * rrobai;
  double rvpfbu = alps * betu * cf * rrobai;
  double rvpfen = alpf * (v2h + db2 + cf * vnbar + cf2 * gammam1i + (cf2 - c2) * gammam2 * gammam1i) - alps * ca * vb * sgn;
  double rvmfro = alpf;
  double rvmfmn = alpf * (vnbar - cf);
  double rvmfmt = alpf * vtbar + alps * bett * ca * sgn;
  double rvmfmu = alpf * vubar + alps * betu * ca * sgn;
  double rvmfbt = rvpfbt;
  double rvmfbu = rvpfbu;
  double rvmfen = alpf * (v2h + db2 - cf * vnbar + cf2 * gammam1i + (cf2 - c2) * gammam2 * gammam1i) + alps * ca * vb * sgn;
  double rvpsro = alps;
  double rvpsmn = alps * (vnbar + cs);
  double rvpsmt = alps * vtbar + c * sgn * alpf * bett;
  double rvpsmu = alps * vubar + c * sgn * alpf * betu;
  double rvpsbt = -rrobai * c2 * alpf * bett * cfi;
  double rvpsbu = -rrobai * c2 * alpf * betu * cfi;
  double rvpsen = alps * (v2h + db2 + cs * vnbar + cs2 * gammam1i + (cs2 - c2) * gammam2 * gammam1i) + alpf * c * vb * sgn;
  double rvmsro = alps;
  double rvmsmn = alps * (vnbar - cs);
  double rvmsmt = alps * vtbar - c * sgn * alpf * bett;
  double rvmsmu = alps * vubar - c * sgn * alpf * betu;
  double rvmsbt = rvpsbt;
  double rvmsbu = rvpsbu;
  double rvmsen = alps * (v2h + db2 - cs * vnbar + cs2 * gammam1i + (cs2 - c2) * gammam2 * gammam1i) - alpf * c * vb * sgn;
  const double rvparo = 0.0, rvpamn = 0.0;
  double rvpamt = -sgn * betu;
  double rvpamu = sgn * bett;
  double rvpabt = rrobai * betu;
  double rvpabu = -rrobai * bett;
  double rvpaen = -(betu * vtbar - bett * vubar) * sgn;
  const double rvmaro = 0.0, rvmamn = 0.0;
  double rvmamt = -rvpamt;
  double rvmamu = -rvpamu;
  double rvmabt = rvpabt;
  double rvmabu = rvpabu;
  double rvmaen = -rvpaen;
  const double rv00ro = 1.0, rv00bt = 0.0, rv00bu = 0.0;
  double rv00mn = vnbar, rv00mt = vtbar, rv00mu = vubar;
  double rv00en = v2h + db2;
#define MLAU_ROE(FL, FR, S)                                                                                   \
  0.5 * (+FL + FR - evpf * wvpf * rvpf##S - evmf * wvmf * rvmf##S - evps * wvps * rvps##S -                   \
         evms * wvms * rvms##S - evpa * wvpa * rvpa##S - evma * wvma * rvma##S - ev00 * wv00 * rv00##S)
  f[0] = MLAU_ROE(frol, fror, ro);
  f[1] = MLAU_ROE(fmnl, fmnr, mn);
  f[2] = MLAU_ROE(fmtl, fmtr, mt);
  f[3] = MLAU_ROE(fmul, fmur, mu);
  f[4] = MLAU_ROE(fbtl, fbtr, bt);
  f[5] = MLAU_ROE(fbul, fbur, bu);
  f[6] = MLAU_ROE(fenl, fenr, en);
#undef MLAU_ROE
}

// include/mhd_flux.h:456-582 (LOWDISS=false) and :584-719 (LOWDISS=true)
template <bool LOWDISS>
MLAU_DEV void flux_hlld(const double (&L)[7], const double (&R)[7], double bnc, double gamma, double dv0,
                        double dv1, double (&f)[7])
{
  const double rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const double ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  const double EPS = 1e-8;
  double gammam1i = 1.0 / (gamma - 1.0);
  double bnc2 = bnc * bnc;
  double sgn = (bnc > 0) ? (1.0) : (-1.0);
  double roli = 1.0 / rol;
  double vl2 = vnl * vnl + vtl * vtl + vul * vul;
  double pml = 0.5 * (btl * btl + bul * bul);
  double ptl = prl + pml;
  double enl = gammam1i * prl + pml + 0.5 * rol * vl2;
  double vbl = vtl * btl + vul * bul;
  double rori = 1.0 / ror;
  double vr2 = vnr * vnr + vtr * vtr + vur * vur;
  double pmr = 0.5 * (btr * btr + bur * bur);
  double ptr = prr + pmr;
  double enr = gammam1i * prr + pmr + 0.5 * ror * vr2;
  double vbr = vtr * btr + vur * bur;
  double cl2 = gamma * prl * roli;
  double cr2 = gamma * prr * rori;
  double cal2 = bnc2 * roli;
  double car2 = bnc2 * rori;
  double cbl2 = cl2 + cal2 + 2.0 * pml * roli;
  double cbr2 = cr2 + car2 + 2.0 * pmr * rori;
  double cfl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * cl2 * cal2)));
  double cfr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * cr2 * car2)));
  double cmax, vnc, ptc, sl, sr, slvl, srvr, rslvl, rsrvr, drsvi;
  if constexpr (LOWDISS) {
    cbl2 = vl2 + cal2 + 2.0 * pml * roli;
    cbr2 = vr2 + car2 + 2.0 * pmr * rori;
    double ccl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * vl2 * cal2)));
    double ccr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * vr2 * car2)));
    cmax = sqrt(dmax(cfl2, cfr2));
    sl = dmin(0.0, dmin(vnl, vnr) - cmax);
    sr = dmax(0.0, dmax(vnl, vnr) + cmax);
    slvl = sl - vnl;
    srvr = sr - vnr;
    rslvl = rol * slvl;
    rsrvr = ror * srvr;
    drsvi = 1.0 / (rsrvr - rslvl);
    double theta = dmin(1.0, (cmax - dmin(dv0, 0.0)) / (cmax - dmin(dv1, 0.0)));
    theta *= theta;
    theta *= theta;
    vnc = (rsrvr * vnr - rslvl * vnl - theta * (ptr - ptl)) * drsvi;
    double mcc = dmin(1.0, sqrt(dmax(ccl2, ccr2)) / cmax);
    ptc = (rsrvr * ptl - rslvl * ptr + (mcc * (2.0 - mcc)) * rsrvr * rslvl * (vnr - vnl)) * drsvi;
  } else {
    (void)dv0;
    (void)dv1;
    cmax = dmax(sqrt(cfl2), sqrt(cfr2));
    sl = dmin(0.0, dmin(vnl, vnr) - cmax);
    sr = dmax(0.0, dmax(vnl, vnr) + cmax);
    slvl = sl - vnl;
    srvr = sr - vnr;
    rslvl = rol * slvl;
    rsrvr = ror * srvr;
    drsvi = 1.0 / (rsrvr - rslvl);
    vnc = (rsrvr * vnr - rslvl * vnl - ptr + ptl) * drsvi;
    ptc = (rsrvr * ptl - rslvl * ptr + rsrvr * rslvl * (vnr - vnl)) * drsvi;
  }
  double slvc = sl - vnc;
  double srvc = sr - vnc;
  double ro2l = rslvl / slvc;
  double ro2r = rsrvr / srvc;
  double rhdl = rslvl * slvc - bnc2;
  double rhdr = rsrvr * srvc - bnc2;
  double vt2l, vu2l, bt2l, bu2l, vt2r, vu2r, bt2r, bu2r;
  if (fabs(rhdl) > EPS) {
    double rhdli = 1.0 / rhdl;
    double rhnvl = (vnl - vnc) * bnc;
    double rhnbl = rslvl * slvl - bnc2;
    vt2l = vtl + rhnvl * rhdli * btl;
    vu2l = vul + rhnvl * rhdli * bul;
    bt2l = rhnbl * rhdli * btl;
    bu2l = rhnbl * rhdli * bul;
  } else {
    vt2l = vtl;
    vu2l = vul;
    bt2l = btl;
    bu2l = bul;
  }
  if (fabs(rhdr) > EPS) {
    double rhdri = 1.0 / rhdr;
    double rhnvr = (vnr - vnc) * bnc;
    double rhnbr = rsrvr * srvr - bnc2;
    vt2r = vtr + rhnvr * rhdri * btr;
    vu2r = vur + rhnvr * rhdri * bur;
    bt2r = rhnbr * rhdri * btr;
    bu2r = rhnbr * rhdri * bur;
  } else {
    vt2r = vtr;
    vu2r = vur;
    bt2r = btr;
    bu2r = bur;
  }
  double vb2l = vt2l * bt2l + vu2l * bu2l;
  double vb2r = vt2r * bt2r + vu2r * bu2r;
  double en2l = (slvl * enl - ptl * vnl + ptc * vnc + bnc * (vbl - vb2l)) / slvc;
  double en2r = (srvr * enr - ptr * vnr + ptc * vnc + bnc * (vbr - vb2r)) / srvc;
  double rro2l = sqrt(ro2l);
  double rro2r = sqrt(ro2r);
  double rro2i = 1.0 / (rro2r + rro2l);
  double vt3m = (rro2r * vt2r + rro2l * vt2l + (bt2r - bt2l) * sgn) * rro2i;
  double vu3m = (rro2r * vu2r + rro2l * vu2l + (bu2r - bu2l) * sgn) * rro2i;
  double bt3m = (rro2l * bt2r + rro2r * bt2l + rro2r * rro2l * (vt2r - vt2l) * sgn) * rro2i;
  double bu3m = (rro2l * bu2r + rro2r * bu2l + rro2r * rro2l * (vu2r - vu2l) * sgn) * rro2i;
  double vb3m = vt3m * bt3m + vu3m * bu3m;
  double en3l = en2l - rro2l * (vb2l - vb3m) * sgn;
  double en3r = en2r + rro2r * (vb2r - vb3m) * sgn;
  double hl = (vnc > 0) ? (1.0) : (0.0);
  double hr = 1.0 - hl;
  double h2l = (vnc - fabs(bnc) / rro2l > 0) ? (1.0) : (0.0);
  double h3l = (1.0 - h2l) * hl;
  double h2r = (vnc + fabs(bnc) / rro2r > 0) ? (0.0) : (1.0);
  double h3r = (1.0 - h2r) * hr;
  double rou = ro2l * hl + ro2r * hr;
  double vtu = vt2l * h2l + vt3m * (h3l + h3r) + vt2r * h2r;
  double vuu = vu2l * h2l + vu3m * (h3l + h3r) + vu2r * h2r;
  double btu = bt2l * h2l + bt3m * (h3l + h3r) + bt2r * h2r;
  double buu = bu2l * h2l + bu3m * (h3l + h3r) + bu2r * h2r;
  double enu = en2l * h2l + en3l * h3l + en3r * h3r + en2r * h2r;
  f[0] = rou * vnc;
  f[1] = rou * vnc * vnc + ptc - bnc2 * 0.5;
  f[2] = rou * vtu * vnc - bnc * btu;
  f[3] = rou * vuu * vnc - bnc * buu;
  f[4] = btu * vnc - bnc * vtu;
  f[5] = buu * vnc - bnc * vuu;
  f[6] = (enu + ptc) * vnc - bnc * (vtu * btu + vuu * buu);
}

template <int FLUX>
MLAU_DEV void riemann(const double (&L)[7], const double (&R)[7], double bn, double gamma, double dv0, double dv1,
                      double (&f)[7])
{
  if constexpr (FLUX == FLUX_MLAU) flux_mlau(L, R, bn, gamma, dv0, dv1, f);
  else if constexpr (FLUX == FLUX_ROE) flux_roe(L, R, bn, gamma, f);
  else if constexpr (FLUX == FLUX_HLLD) flux_hlld<false>(L, R, bn, gamma, dv0, dv1, f);
  else flux_hlld<true>(L, R, bn, gamma, dv0, dv1, f);
}

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
