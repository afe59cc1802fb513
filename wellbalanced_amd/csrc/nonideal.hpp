// The non-ideal face terms of MHD2D in FP64: the resistive terms
// (carafe kernels/mhd_resistive/mhd_resistive_pyccel.py:21-67) and the anisotropic
// heat flux (kernels/mhd_conductive/mhd_conductive_pyccel.py:25-100), per face, from
// the values of the two cells beside it.  Every expression keeps the reference's
// tree; the translation units that use them for the parity path are built with
// -ffp-contract=off.  The fused stage kernels take the cell values from LDS rows,
// the refluxing and the per-kernel twins from the state or prim in HBM.
#pragma once

#include <hip/hip_runtime.h>

namespace wb {

// The uniform coefficients: resistivity eta, conductivities along / across the field.
struct Diffusion {
  double eta, kpar, kperp;
  __device__ __host__ bool conduction() const { return kpar > 0.0 || kperp > 0.0; }
};

// The smallest normal double: the least length of the field in the unit vector.
constexpr double NI_TINY = 2.2250738585072014e-308;

// Monotonized-central limiter of two differences (mhd_conductive_pyccel.py:25-38).
__device__ __forceinline__ double mc_limit(double a, double b) {
  double slope = 0.0;
  if (a * b > 0.0) {
    const double central = 0.5 * (a + b);
    const double limit = __builtin_fabs(a) < __builtin_fabs(b) ? 2.0 * a : 2.0 * b;
    const double ac = __builtin_fabs(central), al = __builtin_fabs(limit);
    const double size = al < ac ? al : ac;
    if (central > 0.0)
      slope = size;
    else if (central < 0.0)
      slope = -size;
  }
  return slope;
}

// The limited gradient of T along a line at a cell, from T at the cells before (tm)
// and after (tp) it (mhd_conductive_pyccel.py:60-64,82-86).
__device__ __forceinline__ double limited_gradient(double tm, double t0, double tp, double h) {
  return mc_limit((tp - t0) / h, (t0 - tm) / h);
}

// The heat flux through one face, subtracted from the energy flux e4
// (mhd_conductive_pyccel.py:65-79): bn the face's normal field; btL/btR and bzL/bzR
// the tangential in-plane field and B_z of the cells below (L) and above (R) the
// face; tL/tR their temperatures p / rho; gL/gR their limited gradients along the
// face; h the spacing across it.
__device__ __forceinline__ void conductive_face(double bn, double btL, double btR, double bzL,
                                                double bzR, double tL, double tR, double gL,
                                                double gR, double h, const Diffusion& d,
                                                double& e4) {
  const double bt = 0.5 * (btL + btR);
  const double bzf = 0.5 * (bzL + bzR);
  double magnitude = __builtin_sqrt(bn * bn + bt * bt + bzf * bzf);
  magnitude = NI_TINY > magnitude ? NI_TINY : magnitude;
  const double b_n = bn / magnitude;
  const double b_t = bt / magnitude;
  const double normal = (tR - tL) / h;
  const double transverse = mc_limit(gL, gR);
  e4 = e4 - (d.kpar * b_n * (b_n * normal + b_t * transverse) + d.kperp * normal);
}

// The resistive terms of an x-face (mhd_resistive_pyccel.py:50-58): cells L = i-1,
// R = i; jz0/jz1 the current at the face's nodes (j, j+1).  b6: the B_z flux.
__device__ __forceinline__ void resistive_xface(double bzL, double bzR, double byL, double byR,
                                                double jz0, double jz1, double dx, double eta,
                                                double& e4, double& b6) {
  const double jy = -(bzR - bzL) / dx;
  const double bz_x = 0.5 * (bzR + bzL);
  const double by_x = 0.5 * (byR + byL);
  const double jz_x = 0.5 * (jz0 + jz1);
  b6 = b6 + eta * jy;
  e4 = e4 + eta * (jy * bz_x - jz_x * by_x);
}

// The resistive terms of a y-face (mhd_resistive_pyccel.py:60-67): cells L = j-1,
// R = j; jz0/jz1 the current at the face's nodes (i, i+1).
__device__ __forceinline__ void resistive_yface(double bzL, double bzR, double bxL, double bxR,
                                                double jz0, double jz1, double dy, double eta,
                                                double& e4, double& b6) {
  const double jx = (bzR - bzL) / dy;
  const double bz_y = 0.5 * (bzR + bzL);
  const double bx_y = 0.5 * (bxR + bxL);
  const double jz_y = 0.5 * (jz0 + jz1);
  b6 = b6 + -eta * jx;
  e4 = e4 + eta * (jz_y * bx_y - jx * bz_y);
}

// J_z at node (I, J) of a haloed block from its face fields (mhd_resistive_pyccel.py:41-46):
// byf(I, J), byf(I-1, J), bxf(I, J), bxf(I, J-1).
__device__ __forceinline__ double node_current(double byf_ij, double byf_im, double bxf_ij,
                                               double bxf_jm, double dx, double dy) {
  return (byf_ij - byf_im) / dx - (bxf_ij - bxf_jm) / dy;
}

// J_z at node (I, J) from A_z on the nodes (row length YN), the face fields taken as
// mhd_primitive_pyccel.py takes them: bxf = dA/dy + b0x, byf = -dA/dx + b0y.
__device__ __forceinline__ double node_current_az(const double* az, int YN, int I, int J,
                                                  double dx, double dy, double b0x,
                                                  double b0y) {
  auto A = [&](int i, int j) { return az[(size_t)i * YN + j]; };
  const double byf_ij = -(A(I + 1, J) - A(I, J)) / dx + b0y;
  const double byf_im = -(A(I, J) - A(I - 1, J)) / dx + b0y;
  const double bxf_ij = (A(I, J + 1) - A(I, J)) / dy + b0x;
  const double bxf_jm = (A(I, J) - A(I, J - 1)) / dy + b0x;
  return node_current(byf_ij, byf_im, bxf_ij, bxf_jm, dx, dy);
}

}  // namespace wb
