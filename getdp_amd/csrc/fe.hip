// Batched Galerkin assembly on the device.
//
// Replaces, for the canonical terms listed in gdp_hip.h, the reference's
// per-element chain Treatment_FemFormulation (Treatment_Formulation.cpp:413)
// -> Cal_GalerkinTermOfFemEquation (Cal_GalerkinTermOfFemEquation.cpp:531)
// -> Cal_AssembleTerm_NoDt -> Dof_AssembleInMat -> LinAlg_AddDoubleInMatrix.
//
// Design: the reference sums every matrix entry in ascending element order
// (elements sorted by number, GeoData.cpp:1018; each element adds its Ek
// before the next element), and within one element term by term.  Here one
// thread owns one matrix row (one unknown DOF): it walks the elements
// incident to the row's entity in ascending element order and, for each
// element and term, computes the row of Ek that belongs to it and adds it into
// the row's CSR slots.  Each entry thus receives the same additions in the
// same order as in the reference, without atomics, and every Ek entry is
// computed with the reference's order of operations (this file is compiled
// without floating-point contraction).  Per-element Jacobians are computed
// once, by a first kernel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "gdp_device.h"

namespace gdp {

namespace {

constexpr int kGeoStride = 20;  // doubles of per-element geometry
constexpr int kMaxTerms = 16;
constexpr int kMaxSpaces = 3;

enum GeoKind { GEO_NONE = 0, GEO_TRI2D = 1, GEO_TET = 2, GEO_TRISUR = 3 };
enum Family { FAM_NODAL = 0, FAM_EDGE = 1 };

struct DevTerm {
  int kind, equ, dof, npoints, ncoef;
  int perp;                     // test and DOF spaces are BF_PerpendicularEdge
  const double *coef;           // device: coef[region * ncoef + c]
  const unsigned char *active;  // device: per region id, term applies
};

struct DevPlan {
  int family;
  int nspaces, nterms, N, nreg;
  const int *etype, *region, *conn, *edges, *geokind;
  const double *geo;            // nelem * kGeoStride
  const int *dofnum[kMaxSpaces];
  const double *fixval[kMaxSpaces];
  const int *inc_ptr, *inc_item;
  const int *row_space, *row_ent;
  const int *ia, *ja;
  DevTerm term[kMaxTerms];
  int nelem, nmax;
  int bf[kMaxSpaces];  // basis function of each space (GDP_BF_*)
  // element pass output: elements in chunks of 64 (one wavefront); value i of
  // element e sits at ek_cb[e / 64] + 64 * i + e % 64, so that a wavefront's
  // stores of value i are contiguous
  const long *ek_cb;
};

// ---- reference element data ---------------------------------------------------
// Gauss tables with the reference's truncated literals (Gauss_Triangle.h:8-22,
// Gauss_Tetrahedron.h:8-23).
__device__ inline void gauss_pt(int tri, int np, int q, double &u, double &v, double &w,
                                double &wt)
{
  w = 0.;
  if(tri) {
    if(np == 1) {
      u = 0.333333333333333;
      v = 0.333333333333333;
      wt = 0.5;
    }
    else if(np == 3) {
      u = (q == 1) ? 0.66666666666666 : 0.16666666666666;
      v = (q == 2) ? 0.66666666666666 : 0.16666666666666;
      wt = 0.16666666666666;
    }
    else {
      u = q == 0 ? 0.333333333333333 : (q == 1 ? 0.6 : 0.2);
      v = q == 0 ? 0.3333333333333333 : (q == 2 ? 0.6 : 0.2);
      wt = q == 0 ? -0.28125 : .260416666666;
    }
  }
  else {
    if(np == 1) {
      u = v = w = 0.25;
      wt = 0.166666666667;
    }
    else {
      const double a = 0.138196601125, b = 0.585410196625;
      u = (q == 3) ? b : a;
      v = (q == 2) ? b : a;
      w = (q == 1) ? b : a;
      wt = 0.0416666666667;
    }
  }
}

// BF_Node (BF_Node.cpp:49-76)
__device__ inline double bf_node(int tri, int i, double u, double v, double w)
{
  if(tri) return i == 0 ? 1. - u - v : (i == 1 ? u : v);
  return i == 0 ? 1. - u - v - w : (i == 1 ? u : (i == 2 ? v : w));
}

// BF_GradNode (BF_Node.cpp:245-320)
__device__ inline void bf_gradnode(int tri, int i, double s[3])
{
  s[0] = s[1] = s[2] = 0.;
  if(i == 0) {
    s[0] = -1.;
    s[1] = -1.;
    s[2] = tri ? 0. : -1.;
  }
  else
    s[i - 1] = 1.;
}

// BF_Edge / BF_CurlEdge on tetrahedra (BF_Edge.cpp:95-131, 452-486), with the
// sign flip of negative edges (BF_Edge.cpp:366-372, 720-726).
__device__ inline void bf_edge(int k, double u, double v, double w, int neg, double s[3])
{
  switch(k) {
  case 0: s[0] = 1. - v - w; s[1] = u; s[2] = u; break;
  case 1: s[0] = v; s[1] = 1. - u - w; s[2] = v; break;
  case 2: s[0] = w; s[1] = w; s[2] = 1. - u - v; break;
  case 3: s[0] = -v; s[1] = u; s[2] = 0.; break;
  case 4: s[0] = -w; s[1] = 0.; s[2] = u; break;
  default: s[0] = 0.; s[1] = -w; s[2] = v; break;
  }
  if(neg) {
    s[0] = -s[0];
    s[1] = -s[1];
    s[2] = -s[2];
  }
}

__device__ inline void bf_curledge(int k, int neg, double s[3])
{
  switch(k) {
  case 0: s[0] = 0.; s[1] = -2.; s[2] = 2.; break;
  case 1: s[0] = 2.; s[1] = 0.; s[2] = -2.; break;
  case 2: s[0] = -2.; s[1] = 2.; s[2] = 0.; break;
  case 3: s[0] = 0.; s[1] = 0.; s[2] = 2.; break;
  case 4: s[0] = 0.; s[1] = -2.; s[2] = 0.; break;
  default: s[0] = 2.; s[1] = 0.; s[2] = 0.; break;
  }
  if(neg) {
    s[0] = -s[0];
    s[1] = -s[1];
    s[2] = -s[2];
  }
}

// Geometry layout per element (kGeoStride doubles):
//   [0..8]   Jac  c11 c12 c13 c21 c22 c23 c31 c32 c33
//   [9]      DetJac
//   [10..18] InvJac, same order
// ChangeOfCoord_Form1 (Get_Geometry.cpp:1895-1904)
__device__ inline void form1(const double *I, const double u[3], double x[3])
{
  x[0] = u[0] * I[0] + u[1] * I[1] + u[2] * I[2];
  x[1] = u[0] * I[3] + u[1] * I[4] + u[2] * I[5];
  x[2] = u[0] * I[6] + u[1] * I[7] + u[2] * I[8];
}

// ChangeOfCoord_Form2 (Get_Geometry.cpp:1906-1921)
__device__ inline void form2(const double *J, double d, const double u[3], double x[3])
{
  x[0] = (u[0] * J[0] + u[1] * J[3] + u[2] * J[6]) / d;
  x[1] = (u[0] * J[1] + u[1] * J[4] + u[2] * J[7]) / d;
  x[2] = (u[0] * J[2] + u[1] * J[5] + u[2] * J[8]) / d;
}

// Cal_ProductValue (Cal_Value.cpp:641-660, 887-897): coefficient x DOF value.
__device__ inline void coef_times(int ncoef, const double *c, const double v[3], double r[3])
{
  if(ncoef == 1) {
    double a0 = c[0];
    r[0] = a0 * v[0];
    r[1] = a0 * v[1];
    r[2] = a0 * v[2];
  }
  else {
    double t0 = c[0] * v[0] + c[1] * v[1] + c[2] * v[2];
    double t1 = c[3] * v[0] + c[4] * v[1] + c[5] * v[2];
    double t2 = c[6] * v[0] + c[7] * v[1] + c[8] * v[2];
    r[0] = t0;
    r[1] = t1;
    r[2] = t2;
  }
}

// Cal_Product123 (Get_Geometry.cpp:1965-1968)
__device__ inline double dot3(const double a[3], const double b[3])
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// ---- kernel 1: per-element Jacobian, determinant, inverse ----------------------
// JacobianVol2D / JacobianVol3D / JacobianSur3D and Get_InverseMatrix
// (Get_Geometry.cpp:1099-1124, 1435-1469, 1629-1680, 1755-1808), for the
// element with nodes cn[0..3], into G (the layout above).
__device__ inline void geo_of(int gk, const int *__restrict__ cn, const double *__restrict__ xyz, double *G)
{
  int tri = (gk != GEO_TET);
  int nn = tri ? 3 : 4;
  double x[4], y[4], z[4];
  for(int i = 0; i < nn; i++) {
    long nd = cn[i] - 1;
    x[i] = xyz[3 * nd];
    y[i] = xyz[3 * nd + 1];
    z[i] = xyz[3 * nd + 2];
  }
  double J[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  double I[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  double det;
  if(gk == GEO_TRI2D) {
    J[8] = 1.;
    for(int i = 0; i < nn; i++) {
      double g[3];
      bf_gradnode(1, i, g);
      J[0] += x[i] * g[0];
      J[3] += x[i] * g[1];
      J[1] += y[i] * g[0];
      J[4] += y[i] * g[1];
    }
    det = J[0] * J[4] - J[1] * J[3];
    I[0] = J[4] * J[8] / det;
    I[3] = -J[3] * J[8] / det;
    I[1] = -J[1] * J[8] / det;
    I[4] = J[0] * J[8] / det;
    I[8] = 1. / J[8];
  }
  else if(gk == GEO_TRISUR) {
    for(int i = 0; i < nn; i++) {
      double g[3];
      bf_gradnode(1, i, g);
      J[0] += x[i] * g[0];
      J[3] += x[i] * g[1];
      J[1] += y[i] * g[0];
      J[4] += y[i] * g[1];
      J[2] += z[i] * g[0];
      J[5] += z[i] * g[1];
    }
    double a1 = J[0] * J[4] - J[1] * J[3];
    double a2 = J[2] * J[3] - J[0] * J[5];
    double a3 = J[1] * J[5] - J[2] * J[4];
    det = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
  }
  else {
    for(int i = 0; i < nn; i++) {
      double g[3];
      bf_gradnode(0, i, g);
      J[0] += x[i] * g[0];
      J[3] += x[i] * g[1];
      J[6] += x[i] * g[2];
      J[1] += y[i] * g[0];
      J[4] += y[i] * g[1];
      J[7] += y[i] * g[2];
      J[2] += z[i] * g[0];
      J[5] += z[i] * g[1];
      J[8] += z[i] * g[2];
    }
    det = J[0] * J[4] * J[8] + J[2] * J[3] * J[7] + J[1] * J[5] * J[6] -
          J[2] * J[4] * J[6] - J[0] * J[5] * J[7] - J[1] * J[3] * J[8];
    I[0] = (J[4] * J[8] - J[5] * J[7]) / det;
    I[3] = -(J[3] * J[8] - J[5] * J[6]) / det;
    I[6] = (J[3] * J[7] - J[4] * J[6]) / det;
    I[1] = -(J[1] * J[8] - J[2] * J[7]) / det;
    I[4] = (J[0] * J[8] - J[2] * J[6]) / det;
    I[7] = -(J[0] * J[7] - J[1] * J[6]) / det;
    I[2] = (J[1] * J[5] - J[2] * J[4]) / det;
    I[5] = -(J[0] * J[5] - J[2] * J[3]) / det;
    I[8] = (J[0] * J[4] - J[1] * J[3]) / det;
  }
  for(int k = 0; k < 9; k++) G[k] = J[k];
  G[9] = det;
  for(int k = 0; k < 9; k++) G[10 + k] = I[k];
}

__global__ void __launch_bounds__(256) k_geometry(int nelem, const int *__restrict__ geokind,
                                                  const int *__restrict__ conn,
                                                  const double *__restrict__ xyz,
                                                  double *__restrict__ geo)
{
  int e = blockIdx.x * 256 + threadIdx.x;
  if(e >= nelem) return;
  int gk = geokind[e];
  if(gk == GEO_NONE) return;
  geo_of(gk, conn + (long)e * 4, xyz, geo + (long)e * kGeoStride);
}


// Add one Ek entry of the row (Dof_AssembleInMat, DofData.cpp:1464-1500):
// unknown column -> matrix slot; fixed column -> b -= value * Ek (Val != 0).

// ---- two-phase assembly ---------------------------------------------------------
// Pass 1, one thread per element: every active term's element-matrix block
// Ek[k][j] (k = local row, j = local column; source terms Ek[k]), values
// numbered in term order and stored with the chunked layout of ek_cb.  Each value is formed with the
// reference's operation order (Cal_GalerkinTermOfFemEquation.cpp:531-1014:
// quadrature sum of (w * |detJ|) * product, the symmetric terms evaluated
// for j <= i and mirrored, 838-841 and 921-924), compiled without FP
// contraction, so it is bit-identical.
// Pass 2, one thread per CSR entry / per row: the entry's contributions
// summed in (element, term) order, the reference's order of additions
// (Dof_AssembleInMat called element after element, term after term).
// A term's coefficients for region reg (1, 3 or 9 values; the rest zero).
__device__ inline void load_coef(const DevTerm &t, int reg, double cf[9])
{
  const double *c = t.coef + (long)reg * t.ncoef;
#pragma unroll
  for(int k = 0; k < 9; k++) cf[k] = k < t.ncoef ? c[k] : 0.;
}

// BF_PerpendicularEdge terms on a triangle in the plane (2D a_z forms).
// Curl-curl: BF_CurlPerpendicularEdge, the reference-coordinate node gradient
// turned to (dv, -du, dw) (BF_Perpendicular.cpp:70-78), through
// ChangeOfCoord_Form2P (Get_Geometry.cpp:1938-1947), Cal_Product123.  Mass and
// source: the value (0, 0, BF_Node) (:33-41) through ChangeOfCoord_Form1P
// (:1931-1936), Cal_Product3 (:1976).  Returns the next output position.
__device__ double *elem_perp(const DevTerm &t, const double *cf, const double *G, double adet,
                             double *out)
{
  if(t.kind == GDP_T_CURL_CURL) {
    double c[3][3];
    for(int j = 0; j < 3; j++) {
      double u[3];
      bf_gradnode(1, j, u);
      const double r0 = u[1], r1 = -u[0];
      c[j][0] = (r0 * G[0] + r1 * G[3]) / G[9];
      c[j][1] = (r0 * G[1] + r1 * G[4]) / G[9];
      c[j][2] = 0.;
    }
    for(int k = 0; k < 3; k++)
      for(int j = 0; j < 3; j++) {
        const int ii = j > k ? j : k, jj = j > k ? k : j;
        double cc[3];
        coef_times(t.ncoef, cf, c[jj], cc);
        const double pr = dot3(c[ii], cc);
        double acc = 0.;
        for(int q = 0; q < t.npoints; q++) {
          double u, v, w, wt;
          gauss_pt(1, t.npoints, q, u, v, w, wt);
          acc += (wt * adet) * pr;
        }
        out[64 * (k * 3 + j)] = acc;
      }
    return out + 64 * 9;
  }
  if(t.kind == GDP_T_EDGE_MASS) {
    for(int k = 0; k < 3; k++)
      for(int j = 0; j < 3; j++) {
        const int ii = j > k ? j : k, jj = j > k ? k : j;
        double acc = 0.;
        for(int q = 0; q < t.npoints; q++) {
          double u, v, w, wt;
          gauss_pt(1, t.npoints, q, u, v, w, wt);
          const double pi = bf_node(1, ii, u, v, w) / G[8], pj = bf_node(1, jj, u, v, w) / G[8];
          acc += (wt * adet) * (pi * (cf[0] * pj));
        }
        out[64 * (k * 3 + j)] = acc;
      }
    return out + 64 * 9;
  }
  for(int k = 0; k < 3; k++) {  // GDP_T_SRC_EDGE
    double acc = 0.;
    for(int q = 0; q < t.npoints; q++) {
      double u, v, w, wt;
      gauss_pt(1, t.npoints, q, u, v, w, wt);
      acc += (wt * adet) * ((bf_node(1, k, u, v, w) / G[8]) * cf[2]);
    }
    out[64 * k] = acc;
  }
  return out + 64 * 3;
}

// One element's values, NN = 3 (triangles) or 4 (tetrahedra) nodes: the
// node loops are unrolled, so the gradients stay in registers with fixed
// indices.
template <int NN>
__device__ inline void elem_nodal(const DevPlan &P, int e, int gk, double *out)
{
  constexpr int tri = NN == 3;
  const int reg = P.region[e];
  const double *G = P.geo + (long)e * kGeoStride;
  const double adet = fabs(G[9]);
  double g[NN][3];
  if(gk != GEO_TRISUR)
#pragma unroll
    for(int j = 0; j < NN; j++) {
      double u[3];
      bf_gradnode(tri, j, u);
      form1(G + 10, u, g[j]);
    }
  for(int ti = 0; ti < P.nterms; ti++) {
    const DevTerm &t = P.term[ti];
    if(!t.active[reg]) continue;
    // the region's coefficients in registers: the stores to the value array
    // would otherwise force a reload of each coefficient for every value
    double cf[9];
    load_coef(t, reg, cf);
    if(t.perp) {
      out = elem_perp(t, cf, G, adet, out);
      continue;
    }
    if(t.kind == GDP_T_GRAD_GRAD) {
      const bool sym = (t.dof == t.equ);
#pragma unroll
      for(int k = 0; k < NN; k++)
#pragma unroll
        for(int j = 0; j < NN; j++) {
          // symmetric terms: Ek[k][j] for j > k is Ek[j][k] (the mirror)
          const double *gi = sym && j > k ? g[j] : g[k];
          const double *gj = sym && j > k ? g[k] : g[j];
          double cg[3];
          coef_times(t.ncoef, cf, gj, cg);
          double pr = dot3(gi, cg);
          double acc = 0.;
          for(int q = 0; q < t.npoints; q++) {
            double u, v, w, wt;
            gauss_pt(tri, t.npoints, q, u, v, w, wt);
            acc += (wt * adet) * pr;
          }
          out[64 * (k * NN + j)] = acc;
        }
      out += 64 * NN * NN;
    }
    else if(t.kind == GDP_T_NODE_MASS) {
#pragma unroll
      for(int k = 0; k < NN; k++)
#pragma unroll
        for(int j = 0; j < NN; j++) {
          const int ii = j > k ? j : k, jj = j > k ? k : j;
          double acc = 0.;
          for(int q = 0; q < t.npoints; q++) {
            double u, v, w, wt;
            gauss_pt(tri, t.npoints, q, u, v, w, wt);
            double pi = bf_node(tri, ii, u, v, w), pj = bf_node(tri, jj, u, v, w);
            acc += (wt * adet) * (pi * (cf[0] * pj));
          }
          out[64 * (k * NN + j)] = acc;
        }
      out += 64 * NN * NN;
    }
    else if(t.kind == GDP_T_SRC_NODE) {
#pragma unroll
      for(int k = 0; k < NN; k++) {
        double acc = 0.;
        for(int q = 0; q < t.npoints; q++) {
          double u, v, w, wt;
          gauss_pt(tri, t.npoints, q, u, v, w, wt);
          acc += (wt * adet) * (bf_node(tri, k, u, v, w) * cf[0]);
        }
        out[64 * k] = acc;
      }
      out += 64 * NN;
    }
  }
}

__global__ void __launch_bounds__(128) k_elem_nodal(DevPlan P, double *__restrict__ ek)
{
  const int e = blockIdx.x * 128 + threadIdx.x;
  if(e >= P.nelem) return;
  const int gk = P.geokind[e];
  if(gk == GEO_NONE) return;
  double *out = ek + P.ek_cb[e >> 6] + (e & 63);
  if(gk == GEO_TET) elem_nodal<4>(P, e, gk, out);
  else elem_nodal<3>(P, e, gk, out);
}

// Element matrices of symmetric edge terms are formed for j <= k and
// mirrored (value (k, j) = value (max, min)), as the reference does; the
// mapped basis functions (curls, and per quadrature point the edge
// functions) are computed once per element rather than once per value.  Every
// value keeps the reference's operations, so the results are the same bits.
__global__ void __launch_bounds__(128) k_elem_edge(DevPlan P, double *__restrict__ ek)
{
  const int e = blockIdx.x * 128 + threadIdx.x;
  if(e >= P.nelem) return;
  if(P.geokind[e] == GEO_NONE) return;
  const int reg = P.region[e];
  const int *ed = P.edges + (long)e * 6;
  const double *G = P.geo + (long)e * kGeoStride;
  const double det = G[9], adet = fabs(det);
  int neg[6];
#pragma unroll
  for(int k = 0; k < 6; k++) neg[k] = ed[k] < 0;
  double *out = ek + P.ek_cb[e >> 6] + (e & 63);
  for(int ti = 0; ti < P.nterms; ti++) {
    const DevTerm &t = P.term[ti];
    if(!t.active[reg]) continue;
    double cf[9];
    load_coef(t, reg, cf);
    if(t.kind == GDP_T_CURL_CURL) {
      double xc[6][3], cxc[6][3];
#pragma unroll
      for(int k = 0; k < 6; k++) {
        double c[3];
        bf_curledge(k, neg[k], c);
        form2(G, det, c, xc[k]);
        coef_times(t.ncoef, cf, xc[k], cxc[k]);
      }
#pragma unroll
      for(int k = 0; k < 6; k++)
#pragma unroll
        for(int j = 0; j <= k; j++) {
          const double pr = dot3(xc[k], cxc[j]);
          double acc = 0.;
          for(int q = 0; q < t.npoints; q++) {
            double u, v, w, wt;
            gauss_pt(0, t.npoints, q, u, v, w, wt);
            acc += (wt * adet) * pr;
          }
          out[64 * (k * 6 + j)] = acc;
          if(j < k) out[64 * (j * 6 + k)] = acc;
        }
      out += 64 * 36;
    }
    else if(t.kind == GDP_T_EDGE_MASS) {
      double acc[21];
#pragma unroll
      for(int p = 0; p < 21; p++) acc[p] = 0.;
      for(int q = 0; q < t.npoints; q++) {
        double u, v, w, wt;
        gauss_pt(0, t.npoints, q, u, v, w, wt);
        double x[6][3], cx[6][3];
#pragma unroll
        for(int k = 0; k < 6; k++) {
          double sk[3];
          bf_edge(k, u, v, w, neg[k], sk);
          form1(G + 10, sk, x[k]);
          coef_times(t.ncoef, cf, x[k], cx[k]);
        }
#pragma unroll
        for(int k = 0, p = 0; k < 6; k++)
#pragma unroll
          for(int j = 0; j <= k; j++, p++) acc[p] += (wt * adet) * dot3(x[k], cx[j]);
      }
#pragma unroll
      for(int k = 0, p = 0; k < 6; k++)
#pragma unroll
        for(int j = 0; j <= k; j++, p++) {
          out[64 * (k * 6 + j)] = acc[p];
          if(j < k) out[64 * (j * 6 + k)] = acc[p];
        }
      out += 64 * 36;
    }
    else if(t.kind == GDP_T_SRC_EDGE) {
      for(int k = 0; k < 6; k++) {
        double acc = 0.;
        for(int q = 0; q < t.npoints; q++) {
          double u, v, w, wt;
          gauss_pt(0, t.npoints, q, u, v, w, wt);
          double si[3], xi[3];
          bf_edge(k, u, v, w, ed[k] < 0, si);
          form1(G + 10, si, xi);
          acc += (wt * adet) * dot3(xi, cf);
        }
        out[64 * k] = acc;
      }
      out += 64 * 6;
    }
  }
}

// a[s] = sum of the entry's contributions in order (the reference adds each
// Ek value to a zeroed matrix: 0 + v1 + v2 + ...)
__global__ void __launch_bounds__(256) k_slot_sum(long nnz, const int *__restrict__ cptr,
                                                  const int *__restrict__ cidx,
                                                  const double *__restrict__ ek, double *__restrict__ a)
{
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if(s >= nnz) return;
  double acc = 0.;
  const int k1 = cptr[s + 1];
  // up to 8 contributions in flight: their indices, then their values, are
  // loaded together, and then added in order
  for(int k = cptr[s]; k < k1; k += 8) {
    int ix[8];
    double v[8];
#pragma unroll
    for(int j = 0; j < 8; j++) ix[j] = k + j < k1 ? cidx[k + j] : -1;
#pragma unroll
    for(int j = 0; j < 8; j++) v[j] = ix[j] >= 0 ? ek[ix[j]] : 0.;
#pragma unroll
    for(int j = 0; j < 8; j++)
      if(ix[j] >= 0) acc += v[j];
  }
  a[s] = acc;
}

// b[r]: fixed-DOF columns and source terms, b -= (value * 1) * Ek
// (Dof_AssembleInMat unknown x fixed branch, DofData.cpp:1489-1500)
__global__ void __launch_bounds__(256) k_rhs_sum(int N, const int *__restrict__ rptr,
                                                 const int *__restrict__ ridx,
                                                 const double *__restrict__ rfv,
                                                 const double *__restrict__ ek, double *__restrict__ b)
{
  const int r = blockIdx.x * 256 + threadIdx.x;
  if(r >= N) return;
  double bacc = 0.;
  const int k1 = rptr[r + 1];
  // 8 contributions in flight at a time, added in order
  for(int k = rptr[r]; k < k1; k += 8) {
    int ix[8];
    double v[8], f[8];
#pragma unroll
    for(int j = 0; j < 8; j++) {
      ix[j] = k + j < k1 ? ridx[k + j] : -1;
      f[j] = k + j < k1 ? rfv[k + j] : 0.;
    }
#pragma unroll
    for(int j = 0; j < 8; j++) v[j] = ix[j] >= 0 ? ek[ix[j]] : 0.;
#pragma unroll
    for(int j = 0; j < 8; j++)
      if(ix[j] >= 0 && v[j] != 0.) bacc += (f[j] * 1.) * (-v[j]);
  }
  b[r] = bacc;
}

// ---- PRE pass on the device ---------------------------------------------------
// The same tables as the host pass (DOF numbering, incidence, CSR pattern,
// two-phase contribution lists), built in parallel: marks, exclusive scans,
// per-entity and per-row threads.  Every list comes out in the host pass's
// order, so the results are identical.

__device__ inline int nodes_of_d(int et) { return et == GDP_TET ? 4 : et == GDP_TRI ? 3 : et == GDP_LINE ? 2 : 0; }
__device__ inline int nb_of_d(const DevPlan &P, int bf, int e)
{
  const int et = P.etype[e];
  if(bf != GDP_BF_EDGE) return nodes_of_d(et);
  return et == GDP_TET ? 6 : et == GDP_TRI ? 3 : et == GDP_LINE ? 1 : 0;
}
__device__ inline int ent_of_d(const DevPlan &P, int bf, int e, int k)
{
  if(bf != GDP_BF_EDGE) return P.conn[(long)e * 4 + k];
  const int v = P.edges[(long)e * 6 + k];
  return v < 0 ? -v : v;
}

struct PrepOut {
  int *dofnum[kMaxSpaces];
};

// (space, entity) pairs touched by active terms: 1 (Pre_TermOfFemEquation.cpp:125-160)
__global__ void __launch_bounds__(256) k_prep_mark(DevPlan P, PrepOut O, int *err)
{
  const int e = blockIdx.x * 256 + threadIdx.x;
  if(e >= P.nelem) return;
  const int reg = P.region[e];
  for(int ti = 0; ti < P.nterms; ti++) {
    const DevTerm &t = P.term[ti];
    if(!t.active[reg]) continue;
    for(int q = 0; q < 2; q++) {
      const int sp = q ? t.dof : t.equ;
      if(sp < 0) continue;
      const int bf = P.bf[sp], nb = nb_of_d(P, bf, e);
      for(int k = 0; k < nb; k++) {
        const int ent = ent_of_d(P, bf, e, k);
        if(ent < 1 || ent > P.nmax) {
          atomicOr(err, 1);
          continue;
        }
        O.dofnum[sp][ent] = 1;
      }
    }
  }
}

// fixed entities of one space: the first occurrence in the list wins, as the
// reference's constraint loop stops at the first match and never re-assigns
// a fixed DOF (Get_ConstraintOfElement.cpp:51-54, DofData.cpp:1137-1151).
// Marks become kFixMark - f, so the largest mark is the smallest f.
constexpr int kFixMark = 0x7fffffff;
__global__ void __launch_bounds__(256) k_prep_fix(int nfix, const int *fixent, int nmax, int *dofnum)
{
  const int f = blockIdx.x * 256 + threadIdx.x;
  if(f >= nfix) return;
  const int ent = fixent[f];
  if(ent >= 1 && ent <= nmax && dofnum[ent] >= 1) atomicMax(dofnum + ent, kFixMark - f);
}

// unknown flags (1 for a mark that stayed 1) for the numbering scan; fixed
// marks kFixMark - f become -(f + 1)
__global__ void __launch_bounds__(256) k_prep_flags(int nmax, int *dofnum, int *flag)
{
  const int ent = blockIdx.x * 256 + threadIdx.x;
  if(ent > nmax + 1) return;
  if(ent > nmax) {
    flag[ent] = 0;
    return;
  }
  const int v = dofnum[ent];
  flag[ent] = v == 1;
  if(v >= 2) dofnum[ent] = -(kFixMark - v + 1);
}

// numbers in entity order after an exclusive scan of the flags; base = the
// unknowns of the spaces numbered before (NumType order, DofData.cpp:36-45)
__global__ void __launch_bounds__(256) k_prep_number(int nmax, int sp, int base, const int *scan, int *dofnum,
                                                     int *row_space, int *row_ent)
{
  const int ent = blockIdx.x * 256 + threadIdx.x;
  if(ent > nmax || dofnum[ent] != 1) return;
  const int r = base + scan[ent];
  dofnum[ent] = r + 1;
  row_space[r] = sp;
  row_ent[r] = ent;
}

__device__ inline bool elem_used(const DevPlan &P, int reg)
{
  for(int ti = 0; ti < P.nterms; ti++)
    if(P.term[ti].active[reg]) return true;
  return false;
}

// incidence entity -> (element, local index): counts, then fill, then each
// entity's list sorted (elements ascending, as the host's element loop)
__global__ void __launch_bounds__(256) k_prep_inc(DevPlan P, int *cnt, int *cur, int *item, bool fill)
{
  const int e = blockIdx.x * 256 + threadIdx.x;
  if(e >= P.nelem || !elem_used(P, P.region[e])) return;
  const int bf = P.bf[0], nb = nb_of_d(P, bf, e);
  for(int k = 0; k < nb; k++) {
    const int ent = ent_of_d(P, bf, e, k);
    if(fill) item[atomicAdd(cur + ent, 1)] = e * 8 + k;
    else atomicAdd(cnt + ent, 1);
  }
}

__global__ void __launch_bounds__(256) k_prep_inc_sort(int nmax, const int *ptr, int *item)
{
  const int ent = blockIdx.x * 256 + threadIdx.x;
  if(ent > nmax) return;
  const int b = ptr[ent], n = ptr[ent + 1] - b;
  int *a = item + b;
  for(int i = 1; i < n; i++) {
    const int v = a[i];
    int j = i - 1;
    while(j >= 0 && a[j] > v) {
      a[j + 1] = a[j];
      j--;
    }
    a[j + 1] = v;
  }
}

// CSR pattern: one thread per row collects the unknown DOFs of the elements
// around its entity (bilinear terms whose test space is the row's) into a
// sorted, duplicate-free list in LDS; pass 1 counts, pass 2 writes ja.
constexpr int kPrepRow = 128;  // longest row the device pass handles
template <bool FILL>
__global__ void __launch_bounds__(64) k_prep_pattern(DevPlan P, int *rowlen, int *ja, int *err)
{
  __shared__ int buf[kPrepRow * 64];
  const int r = blockIdx.x * 64 + threadIdx.x;
  if(r >= P.N) return;
  int *b = buf + threadIdx.x;
  const int sp = P.row_space[r], er = P.row_ent[r];
  int len = 0;
  for(int p = P.inc_ptr[er]; p < P.inc_ptr[er + 1]; p++) {
    const int e = P.inc_item[p] >> 3, reg = P.region[e];
    for(int ti = 0; ti < P.nterms; ti++) {
      const DevTerm &t = P.term[ti];
      if(t.dof < 0 || t.equ != sp || !t.active[reg]) continue;
      const int bf = P.bf[t.dof], nb = nb_of_d(P, bf, e);
      for(int j = 0; j < nb; j++) {
        const int c = P.dofnum[t.dof][ent_of_d(P, bf, e, j)] - 1;
        if(c < 0) continue;
        int lo = 0, hi = len;  // first position with b[pos] >= c
        while(lo < hi) {
          const int mid = (lo + hi) >> 1;
          if(b[mid * 64] < c) lo = mid + 1;
          else hi = mid;
        }
        if(lo < len && b[lo * 64] == c) continue;
        if(len == kPrepRow) {
          atomicOr(err, 2);
          continue;
        }
        for(int q = len; q > lo; q--) b[q * 64] = b[(q - 1) * 64];
        b[lo * 64] = c;
        len++;
      }
    }
  }
  if(!FILL) {
    rowlen[r] = len;
    if(len == 0) atomicMin(err + 1, r);
    return;
  }
  const int o = P.ia[r];
  for(int i = 0; i < len; i++) ja[o + i] = b[i * 64];
}

// values per element in the two-phase layout; chunk widths of 64 elements
__device__ inline bool bilinear_d(int kind)
{
  return kind == GDP_T_GRAD_GRAD || kind == GDP_T_NODE_MASS || kind == GDP_T_CURL_CURL ||
         kind == GDP_T_EDGE_MASS;
}
__global__ void __launch_bounds__(64) k_prep_chunks(DevPlan P, int *cw)
{
  const int e = blockIdx.x * 64 + threadIdx.x;
  int sz = 0;
  if(e < P.nelem && P.geokind[e] != GEO_NONE) {
    const int nb = P.family == FAM_EDGE ? 6 : nodes_of_d(P.etype[e]), reg = P.region[e];
    for(int ti = 0; ti < P.nterms; ti++)
      if(P.term[ti].active[reg]) sz += bilinear_d(P.term[ti].kind) ? nb * nb : nb;
  }
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) sz = max(sz, __shfl_xor(sz, o, 64));
  if(threadIdx.x == 0) cw[blockIdx.x] = 64 * sz;
}

__global__ void __launch_bounds__(256) k_prep_widen(int n, const int *a, long *b)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if(i <= n) b[i] = a[i];
}

// One row's contributions in (element, term, column) order: matrix entries
// (count per CSR entry, then the value indices) and right-hand side (fixed
// columns with their value, sources with 1).
template <bool FILL>
__global__ void __launch_bounds__(256) k_prep_visit(DevPlan P, const int *cb, int *ccnt, int *cur, int *cidx,
                                                    int *rcnt, const int *rptr, int *ridx, double *rfv)
{
  const int r = blockIdx.x * 256 + threadIdx.x;
  if(r >= P.N) return;
  const bool edge = P.family == FAM_EDGE;
  const int sp = P.row_space[r], er = P.row_ent[r];
  const int rs = P.ia[r], re = P.ia[r + 1];
  int rc = FILL ? rptr[r] : 0;
  for(int p = P.inc_ptr[er]; p < P.inc_ptr[er + 1]; p++) {
    const int item = P.inc_item[p];
    const int e = item >> 3, k = item & 7;
    const int reg = P.region[e], nb = edge ? 6 : nodes_of_d(P.etype[e]);
    const int base = cb[e >> 6] + (e & 63);
    int o = 0;
    for(int ti = 0; ti < P.nterms; ti++) {
      const DevTerm &t = P.term[ti];
      if(!t.active[reg]) continue;
      const bool bil = bilinear_d(t.kind);
      if(edge || t.equ == sp) {
        if(bil)
          for(int j = 0; j < nb; j++) {
            const int ej = edge ? abs(P.edges[(long)e * 6 + j]) : P.conn[(long)e * 4 + j];
            const int c = P.dofnum[t.dof][ej];
            const int vi = base + 64 * (o + k * nb + j);
            if(c > 0) {
              int lo = rs, hi = re;
              while(lo < hi) {
                const int mid = (lo + hi) >> 1;
                if(P.ja[mid] < c - 1) lo = mid + 1;
                else hi = mid;
              }
              if(FILL) cidx[cur[lo]++] = vi;
              else ccnt[lo]++;
            }
            else if(c < 0) {
              if(FILL) {
                ridx[rc] = vi;
                rfv[rc] = P.fixval[t.dof][-c - 1];
              }
              rc++;
            }
          }
        else {
          if(FILL) {
            ridx[rc] = base + 64 * (o + k);
            rfv[rc] = 1.0;
          }
          rc++;
        }
      }
      o += bil ? nb * nb : nb;
    }
  }
  if(!FILL) rcnt[r] = rc;
}

// ---- exclusive scan of int arrays (in place, deterministic) ----
__global__ void __launch_bounds__(256) k_scan_local(int *d, long n, int *bsum)
{
  __shared__ int sh[256];
  const long base = (long)blockIdx.x * 2048 + threadIdx.x * 8;
  int v[8], t = 0;
#pragma unroll
  for(int i = 0; i < 8; i++) {
    v[i] = base + i < n ? d[base + i] : 0;
    t += v[i];
  }
  sh[threadIdx.x] = t;
  __syncthreads();
  for(int o = 1; o < 256; o <<= 1) {
    const int x = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += x;
    __syncthreads();
  }
  int run = threadIdx.x ? sh[threadIdx.x - 1] : 0;
#pragma unroll
  for(int i = 0; i < 8; i++) {
    if(base + i < n) d[base + i] = run;
    run += v[i];
  }
  if(threadIdx.x == 255) bsum[blockIdx.x] = sh[255];
}

__global__ void __launch_bounds__(256) k_scan_add(int *d, long n, const int *boff)
{
  const long base = (long)blockIdx.x * 2048 + threadIdx.x * 8;
  const int add = boff[blockIdx.x];
#pragma unroll
  for(int i = 0; i < 8; i++)
    if(base + i < n) d[base + i] += add;
}

int nodes_of(int et) { return et == GDP_TET ? 4 : et == GDP_TRI ? 3 : et == GDP_LINE ? 2 : 0; }
int edges_of(int et) { return et == GDP_TET ? 6 : et == GDP_TRI ? 3 : et == GDP_LINE ? 1 : 0; }

}  // namespace
}  // namespace gdp

using namespace gdp;

// =============================================================================
// FE system object
// =============================================================================
struct gdp_fe {
  int family = FAM_NODAL;
  int nelem = 0, nnodes = 0, nedges = 0, nmax = 0, nreg = 0;
  int nspaces = 0, nterms = 0, N = 0;
  long nnz = 0;
  // host copies
  std::vector<int> h_etype, h_region, h_conn, h_edges, h_geokind;
  std::vector<int> h_perp;  // per term: BF_PerpendicularEdge spaces
  int h_bf[kMaxSpaces] = {0, 0, 0}, h_numtype[kMaxSpaces] = {0, 0, 0};
  std::vector<std::vector<int>> h_fixent;  // fixed entities per space
  std::vector<std::vector<int>> h_dofnum;  // per space, nmax+1
  std::vector<int> h_ia, h_ja;
  std::vector<int> h_inc_ptr, h_inc_item, h_row_space, h_row_ent;
  std::vector<double> h_xyz;
  std::vector<std::vector<double>> h_fixval, h_coef;
  std::vector<std::vector<unsigned char>> h_active;
  std::vector<std::vector<unsigned char>> h_support;  // per space, per region
  std::vector<gdp_term_desc> h_terms;
  bool uploaded = false;
  bool prepped_on_device = false;
  // device
  int *d_etype = nullptr, *d_region = nullptr, *d_conn = nullptr, *d_edges = nullptr,
      *d_geokind = nullptr;
  double *d_xyz = nullptr, *d_geo = nullptr;
  int *d_dofnum[kMaxSpaces] = {nullptr, nullptr, nullptr};
  double *d_fixval[kMaxSpaces] = {nullptr, nullptr, nullptr};
  int *d_inc_ptr = nullptr, *d_inc_item = nullptr, *d_row_space = nullptr, *d_row_ent = nullptr;
  int *d_ia = nullptr, *d_ja = nullptr;
  double *d_coef[kMaxTerms] = {};
  unsigned char *d_active[kMaxTerms] = {};
  DevPlan plan{};
  // two-phase assembly tables (built on the first Generate)
  long ek_total = 0;
  long *d_ek_cb = nullptr;
  double *d_ek = nullptr, *d_rfv = nullptr;
  int *d_cptr = nullptr, *d_cidx = nullptr, *d_rptr = nullptr, *d_ridx = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  double ms_total = 0, ms_geo = 0, ms_gather = 0;

  ~gdp_fe() { release_device(); }
  void release_device()
  {
    dfree(d_etype);
    dfree(d_region);
    dfree(d_conn);
    dfree(d_edges);
    dfree(d_geokind);
    dfree(d_xyz);
    dfree(d_geo);
    for(int s = 0; s < kMaxSpaces; s++) {
      dfree(d_dofnum[s]);
      dfree(d_fixval[s]);
    }
    dfree(d_inc_ptr);
    dfree(d_inc_item);
    dfree(d_row_space);
    dfree(d_row_ent);
    dfree(d_ia);
    dfree(d_ja);
    for(int t = 0; t < kMaxTerms; t++) {
      dfree(d_coef[t]);
      dfree(d_active[t]);
    }
    dfree(d_ek_cb);
    dfree(d_ek);
    dfree(d_rfv);
    dfree(d_cptr);
    dfree(d_cidx);
    dfree(d_rptr);
    dfree(d_ridx);
    for(auto &e : ev) {
      if(e) (void)hipEventDestroy(e);
      e = nullptr;
    }
    plan = DevPlan{};
    uploaded = false;
  }
};

static bool fe_prep_on_device();
static int fe_prep_device(gdp_fe *F);
static void fe_free_device(gdp_fe *F) { F->release_device(); }
constexpr int kPrepFallback = 1;  // the device PRE pass hands over to the host pass

namespace {

bool contains(int n, const int *l, int v)
{
  for(int i = 0; i < n; i++)
    if(l[i] == v) return true;
  return false;
}

// Entity k of element e in a space: node number or |edge number|.
inline int entity_of(const gdp_fe &F, int bf, int e, int k)
{
  if(bf != GDP_BF_EDGE) return F.h_conn[(size_t)e * 4 + k];
  int v = F.h_edges[(size_t)e * 6 + k];
  return v < 0 ? -v : v;
}

template <typename Fn> void parallel_for(int n, Fn fn)
{
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if(n < 4096 || nt == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  int chunk = (n + (int)nt - 1) / (int)nt;
  for(unsigned t = 0; t < nt; t++) {
    int lo = (int)t * chunk, hi = std::min(n, lo + chunk);
    if(lo >= hi) break;
    th.emplace_back([=]() { fn(lo, hi); });
  }
  for(auto &x : th) x.join();
}

}  // namespace

// PRE pass on the host: DOF numbering (DofData.cpp:36-45, 1279-1323),
// entity -> element incidence, CSR pattern.
static int fe_prep_host(gdp_fe *F)
{
  const int E = F->nelem, nspaces = F->nspaces, nterms = F->nterms;
  const int nmax = F->nmax;
  F->h_dofnum.assign(nspaces, std::vector<int>((size_t)nmax + 1, 0));
  for(int e = 0; e < E; e++) {
    int reg = F->h_region[e];
    for(int ti = 0; ti < nterms; ti++) {
      if(!F->h_active[ti][reg]) continue;
      int sps[2] = {F->h_terms[ti].equ, F->h_terms[ti].dof};
      for(int q = 0; q < 2; q++) {
        int s = sps[q];
        if(s < 0) continue;
        int nb = F->h_bf[s] != GDP_BF_EDGE ? nodes_of(F->h_etype[e]) : edges_of(F->h_etype[e]);
        for(int k = 0; k < nb; k++) {
          int ent = entity_of(*F, F->h_bf[s], e, k);
          if(ent < 1 || ent > nmax) return set_error(GDP_E_ARG, "entity %d out of range", ent);
          F->h_dofnum[s][ent] = 1;
        }
      }
    }
  }
  for(int s = 0; s < nspaces; s++)
    for(size_t f = 0; f < F->h_fixent[s].size(); f++) {
      int ent = F->h_fixent[s][f];
      // first occurrence wins: only an unknown mark becomes fixed
      // (Get_ConstraintOfElement.cpp:51-54, DofData.cpp:1137-1151)
      if(ent >= 1 && ent <= nmax && F->h_dofnum[s][ent] > 0) F->h_dofnum[s][ent] = -(f + 1);
    }
  std::vector<int> order(nspaces);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return F->h_numtype[x] < F->h_numtype[y]; });
  std::vector<int> row_space, row_ent;
  int N = 0;
  for(int s : order)
    for(int ent = 1; ent <= nmax; ent++)
      if(F->h_dofnum[s][ent] > 0) {
        F->h_dofnum[s][ent] = ++N;
        row_space.push_back(s);
        row_ent.push_back(ent);
      }
  F->N = N;

  // ---- incidence: entity -> (element, local index), elements ascending ----
  int space_bf = F->h_bf[0];
  std::vector<int> inc_ptr((size_t)nmax + 2, 0), inc_item;
  auto elem_used = [&](int e) {
    int reg = F->h_region[e];
    for(int ti = 0; ti < nterms; ti++)
      if(F->h_active[ti][reg]) return true;
    return false;
  };
  for(int e = 0; e < E; e++) {
    if(!elem_used(e)) continue;
    int nb = space_bf != GDP_BF_EDGE ? nodes_of(F->h_etype[e]) : edges_of(F->h_etype[e]);
    for(int k = 0; k < nb; k++) inc_ptr[entity_of(*F, space_bf, e, k) + 1]++;
  }
  for(int i = 0; i <= nmax; i++) inc_ptr[i + 1] += inc_ptr[i];
  inc_item.resize(inc_ptr[nmax + 1]);
  {
    std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for(int e = 0; e < E; e++) {
      if(!elem_used(e)) continue;
      int nb = space_bf != GDP_BF_EDGE ? nodes_of(F->h_etype[e]) : edges_of(F->h_etype[e]);
      for(int k = 0; k < nb; k++) inc_item[fill[entity_of(*F, space_bf, e, k)]++] = e * 8 + k;
    }
  }

  // ---- CSR pattern: all unknown DOF pairs sharing an element in a bilinear
  // term, stored zeros included (F6; csr_format sorts columns) ----
  std::vector<int> rowlen(N, 0);
  std::vector<std::vector<int>> rowcols(N);
  parallel_for(N, [&](int lo, int hi) {
    std::vector<int> cand;
    for(int r = lo; r < hi; r++) {
      int s = row_space[r], ent = row_ent[r];
      cand.clear();
      for(int p = inc_ptr[ent]; p < inc_ptr[ent + 1]; p++) {
        int e = inc_item[p] >> 3;
        int reg = F->h_region[e];
        for(int ti = 0; ti < nterms; ti++) {
          const gdp_term_desc &t = F->h_terms[ti];
          if(t.dof < 0 || t.equ != s || !F->h_active[ti][reg]) continue;
          int nb = F->h_bf[t.dof] != GDP_BF_EDGE ? nodes_of(F->h_etype[e]) : edges_of(F->h_etype[e]);
          for(int j = 0; j < nb; j++) {
            int c = F->h_dofnum[t.dof][entity_of(*F, F->h_bf[t.dof], e, j)];
            if(c > 0) cand.push_back(c - 1);
          }
        }
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      rowcols[r] = cand;
      rowlen[r] = (int)cand.size();
    }
  });
  F->h_ia.assign((size_t)N + 1, 0);
  for(int r = 0; r < N; r++) {
    if(!rowlen[r]) return set_error(GDP_E_ZERO_ROW, "Zero row in matrix (DOF %d)", r + 1);
    F->h_ia[r + 1] = F->h_ia[r] + rowlen[r];
  }
  F->nnz = F->h_ia[N];
  if(F->nnz > 2147483647L) return set_error(GDP_E_UNSUPPORTED, "nnz exceeds int32");
  F->h_ja.resize((size_t)F->nnz);
  parallel_for(N, [&](int lo, int hi) {
    for(int r = lo; r < hi; r++)
      std::copy(rowcols[r].begin(), rowcols[r].end(), F->h_ja.begin() + F->h_ia[r]);
  });
  rowcols.clear();
  rowcols.shrink_to_fit();

  // ---- keep what the device upload needs (done on first Generate) ----
  F->h_inc_ptr.swap(inc_ptr);
  F->h_inc_item.swap(inc_item);
  F->h_row_space.swap(row_space);
  F->h_row_ent.swap(row_ent);
  return GDP_OK;
}

extern "C" int gdp_fe_create(const gdp_mesh_desc *mesh, int nspaces,
                             const gdp_space_desc *spaces, int nterms,
                             const gdp_term_desc *terms, gdp_fe **out)
{
  if(!mesh || !spaces || !terms || !out || nspaces < 1 || nterms < 1)
    return set_error(GDP_E_ARG, "gdp_fe_create: bad arguments");
  if(nspaces > kMaxSpaces || nterms > kMaxTerms)
    return set_error(GDP_E_UNSUPPORTED, "at most %d spaces and %d terms", kMaxSpaces, kMaxTerms);
  std::unique_ptr<gdp_fe> F(new gdp_fe);
  F->nelem = mesh->nelem;
  F->nnodes = mesh->nnodes;
  F->nedges = mesh->edges ? mesh->nedges : 0;
  F->nmax = std::max(F->nnodes, F->nedges);
  F->nspaces = nspaces;
  F->nterms = nterms;
  const int E = F->nelem;
  F->h_etype.assign(mesh->etype, mesh->etype + E);
  F->h_region.assign(mesh->region, mesh->region + E);
  F->h_conn.assign(mesh->conn, mesh->conn + (size_t)E * 4);
  if(mesh->edges) F->h_edges.assign(mesh->edges, mesh->edges + (size_t)E * 6);
  int maxreg = 0;
  for(int e = 0; e < E; e++) {
    if(F->h_region[e] < 0) return set_error(GDP_E_ARG, "negative region id");
    maxreg = std::max(maxreg, F->h_region[e]);
  }
  F->nreg = maxreg + 1;

  // ---- classification (Cal_InitGalerkinTermOfFemEquation analogue) ----
  int nedge_sp = 0;
  for(int s = 0; s < nspaces; s++) {
    if(spaces[s].bf == GDP_BF_EDGE) nedge_sp++;
    else if(spaces[s].bf != GDP_BF_NODE && spaces[s].bf != GDP_BF_PERP_EDGE)
      return set_error(GDP_E_UNSUPPORTED, "space %d: unknown basis function", s);
  }
  if(nedge_sp && (nspaces != 1 || !mesh->edges))
    return set_error(GDP_E_UNSUPPORTED, "edge spaces: exactly one space and mesh edges needed");
  F->family = nedge_sp ? FAM_EDGE : FAM_NODAL;
  F->h_geokind.assign(E, GEO_NONE);
  for(int ti = 0; ti < nterms; ti++) {
    const gdp_term_desc &t = terms[ti];
    if(t.equ < 0 || t.equ >= nspaces || t.dof < -1 || t.dof >= nspaces)
      return set_error(GDP_E_ARG, "term %d: bad space index", ti);
    bool src = (t.kind == GDP_T_SRC_NODE || t.kind == GDP_T_SRC_EDGE);
    if(src != (t.dof < 0)) return set_error(GDP_E_ARG, "term %d: source terms have no Dof", ti);
    bool edgek = (t.kind == GDP_T_CURL_CURL || t.kind == GDP_T_EDGE_MASS || t.kind == GDP_T_SRC_EDGE);
    bool nodek = (t.kind == GDP_T_GRAD_GRAD || t.kind == GDP_T_NODE_MASS || t.kind == GDP_T_SRC_NODE);
    if(!edgek && !nodek) return set_error(GDP_E_UNSUPPORTED, "term %d: kind %d", ti, t.kind);
    // BF_PerpendicularEdge lives on nodes: the edge kinds in their 2D
    // perpendicular forms (Form1P / Form2P), on both sides of the term
    const bool perp = spaces[t.equ].bf == GDP_BF_PERP_EDGE;
    if(perp && (!edgek || (t.dof >= 0 && spaces[t.dof].bf != GDP_BF_PERP_EDGE)))
      return set_error(GDP_E_UNSUPPORTED, "term %d: kind %d on a perpendicular-edge space", ti, t.kind);
    if(!perp && t.dof >= 0 && spaces[t.dof].bf == GDP_BF_PERP_EDGE)
      return set_error(GDP_E_UNSUPPORTED, "term %d: mixed forms", ti);
    if(!perp && edgek != (F->family == FAM_EDGE))
      return set_error(GDP_E_UNSUPPORTED, "term %d: mixes edge and nodal spaces", ti);
    if(perp && t.jac_sur)
      return set_error(GDP_E_UNSUPPORTED, "term %d: perpendicular edges need Jacobian Vol in 2D", ti);
    F->h_perp.push_back(perp);
    int need_coef = (t.kind == GDP_T_GRAD_GRAD || t.kind == GDP_T_CURL_CURL) ? -1
                    : (t.kind == GDP_T_SRC_EDGE ? 3 : 1);
    if(need_coef == -1 ? (t.ncoef != 1 && t.ncoef != 9) : t.ncoef != need_coef)
      return set_error(GDP_E_UNSUPPORTED, "term %d: %d coefficient values", ti, t.ncoef);
    if(t.coef_nreg < F->nreg) return set_error(GDP_E_ARG, "term %d: coefficient table too short", ti);
    if(!t.coef || !t.regions) return set_error(GDP_E_ARG, "term %d: null table", ti);
    std::vector<unsigned char> inreg(F->nreg, 0);  // region r in the term's regions
    for(int q = 0; q < t.nregions; q++)
      if(t.regions[q] >= 0 && t.regions[q] < F->nreg) inreg[t.regions[q]] = 1;
    for(int e = 0; e < E; e++) {
      if(!inreg[F->h_region[e]]) continue;
      int et = F->h_etype[e], gk;
      if(et == GDP_TET) {
        if(perp) return set_error(GDP_E_UNSUPPORTED, "term %d: perpendicular edges on a tetrahedron", ti);
        if(t.jac_sur) return set_error(GDP_E_UNSUPPORTED, "term %d: Jacobian Sur on a tetrahedron", ti);
        gk = GEO_TET;
        if(t.npoints != 1 && t.npoints != 4)
          return set_error(GDP_E_UNSUPPORTED, "term %d: %d Gauss points on tetrahedra", ti, t.npoints);
      }
      else if(et == GDP_TRI) {
        if(edgek && !perp) return set_error(GDP_E_UNSUPPORTED, "term %d: edge elements on triangles", ti);
        gk = t.jac_sur ? GEO_TRISUR : GEO_TRI2D;
        if(t.jac_sur && t.kind != GDP_T_SRC_NODE)
          return set_error(GDP_E_UNSUPPORTED, "term %d: Jacobian Sur only for sources", ti);
        if(t.npoints != 1 && t.npoints != 3 && t.npoints != 4)
          return set_error(GDP_E_UNSUPPORTED, "term %d: %d Gauss points on triangles", ti, t.npoints);
      }
      else
        return set_error(GDP_E_UNSUPPORTED, "term %d: element type %d", ti, et);
      if(F->h_geokind[e] != GEO_NONE && F->h_geokind[e] != gk)
        return set_error(GDP_E_UNSUPPORTED, "element %d used with two Jacobian types", e + 1);
      F->h_geokind[e] = gk;
    }
  }
  for(int e = 0; e < E; e++)
    for(int k = 0; k < nodes_of(F->h_etype[e]); k++) {
      int nd = F->h_conn[(size_t)e * 4 + k];
      if(F->h_geokind[e] != GEO_NONE && (nd < 1 || nd > F->nnodes))
        return set_error(GDP_E_ARG, "element %d: node %d out of range", e + 1, nd);
    }

  // term activity per region: the element's region is in the term's regions
  // and in the supports of the spaces the term uses (Get_DofOfElement.cpp:102-110)
  std::vector<std::vector<unsigned char>> active(nterms, std::vector<unsigned char>(F->nreg, 0));
  for(int ti = 0; ti < nterms; ti++) {
    const gdp_term_desc &t = terms[ti];
    for(int r = 0; r < F->nreg; r++) {
      bool on = contains(t.nregions, t.regions, r) &&
                contains(spaces[t.equ].nsupport, spaces[t.equ].support, r) &&
                (t.dof < 0 || contains(spaces[t.dof].nsupport, spaces[t.dof].support, r));
      active[ti][r] = on;
    }
  }

  // ---- the formulation's tables (both pre-processing paths use them) ----
  F->h_xyz.assign(mesh->xyz, mesh->xyz + (size_t)F->nnodes * 3);
  F->h_fixval.resize(nspaces);
  F->h_fixent.resize(nspaces);
  for(int s = 0; s < nspaces; s++) {
    F->h_fixval[s].assign(std::max(1, spaces[s].nfixed), 0.);
    F->h_fixent[s].assign(spaces[s].fixed_entity, spaces[s].fixed_entity + spaces[s].nfixed);
    for(int f = 0; f < spaces[s].nfixed; f++) F->h_fixval[s][f] = spaces[s].fixed_value[f];
    F->h_bf[s] = spaces[s].bf;
    F->h_numtype[s] = spaces[s].numtype;
  }
  F->h_active.swap(active);
  F->h_support.assign(nspaces, std::vector<unsigned char>(F->nreg, 0));
  for(int sp = 0; sp < nspaces; sp++)
    for(int q = 0; q < spaces[sp].nsupport; q++)
      if(spaces[sp].support[q] >= 0 && spaces[sp].support[q] < F->nreg) F->h_support[sp][spaces[sp].support[q]] = 1;
  F->h_coef.resize(nterms);
  F->h_terms.assign(terms, terms + nterms);
  for(int ti = 0; ti < nterms; ti++) {
    F->h_coef[ti].assign(terms[ti].coef, terms[ti].coef + (size_t)F->nreg * terms[ti].ncoef);
    F->h_terms[ti].coef = nullptr;
    F->h_terms[ti].regions = nullptr;
  }
  // ---- PRE pass: on the device when the library is initialised ----
  if(fe_prep_on_device()) {
    const int rc = fe_prep_device(F.get());
    if(rc == GDP_OK) {
      *out = F.release();
      return GDP_OK;
    }
    if(rc != kPrepFallback) return rc;
    fe_free_device(F.get());
  }
  GDP_TRY(fe_prep_host(F.get()));
  *out = F.release();
  return GDP_OK;
}

// Tables of the two-phase assembly.  Element values: for each element, its
// active terms in order, nb x nb values (bilinear terms) or nb values
// (sources), in the chunked layout (value i of element e at
// cb[e / 64] + 64 i + e % 64).  For each CSR entry, the block values that add into it, in
// (element, term) order; for each row, the values that add into b (fixed
// columns with their fixed value, sources with 1), in (element, term, column)
// order -- the order of the reference's element loop.
static int fe_build_two_phase(gdp_fe *F)
{
  const int E = F->nelem, N = F->N, nterms = F->nterms;
  const bool edge = F->family == FAM_EDGE;
  auto nb_of = [&](int e) { return edge ? 6 : nodes_of(F->h_etype[e]); };
  auto bilinear = [](int kind) {
    return kind == GDP_T_GRAD_GRAD || kind == GDP_T_NODE_MASS || kind == GDP_T_CURL_CURL ||
           kind == GDP_T_EDGE_MASS;
  };
  // values per element, then chunk bases: a chunk of 64 elements is as wide
  // as its widest element
  const int nch = (E + 63) / 64;
  std::vector<long> cb((size_t)nch + 1, 0);
  for(int c = 0; c < nch; c++) {
    long w = 0;
    for(int e = c * 64; e < std::min(E, c * 64 + 64); e++) {
      long sz = 0;
      if(F->h_geokind[e] != GEO_NONE) {
        const int nb = nb_of(e), reg = F->h_region[e];
        for(int ti = 0; ti < nterms; ti++)
          if(F->h_active[ti][reg]) sz += bilinear(F->h_terms[ti].kind) ? nb * nb : nb;
      }
      w = std::max(w, sz);
    }
    cb[c + 1] = cb[c] + 64 * w;
  }
  F->ek_total = cb[nch];
  if(cb[nch] > 2147483647L) return set_error(GDP_E_UNSUPPORTED, "element blocks exceed int32 indexing");
  auto vidx = [&](int e, long i) { return (int)(cb[e >> 6] + 64 * i + (e & 63)); };
  const int *ia = F->h_ia.data(), *ja = F->h_ja.data();
  // visit(r, fn_slot, fn_rhs): the row's contributions in order
  auto visit = [&](int r, auto &&on_slot, auto &&on_rhs) {
    const int s = F->h_row_space[r], ent = F->h_row_ent[r];
    const int rs = ia[r], re = ia[r + 1];
    for(int p = F->h_inc_ptr[ent]; p < F->h_inc_ptr[ent + 1]; p++) {
      const int item = F->h_inc_item[p];
      const int e = item >> 3, k = item & 7;
      const int reg = F->h_region[e], nb = nb_of(e);
      long o = 0;
      for(int ti = 0; ti < nterms; ti++) {
        if(!F->h_active[ti][reg]) continue;
        const gdp_term_desc &t = F->h_terms[ti];
        const bool bil = bilinear(t.kind);
        if(edge || t.equ == s) {
          if(bil)
            for(int j = 0; j < nb; j++) {
              const int ej = edge ? std::abs(F->h_edges[(size_t)e * 6 + j]) : F->h_conn[(size_t)e * 4 + j];
              const int c = F->h_dofnum[t.dof][ej];
              if(c > 0) {
                const int slot = (int)(std::lower_bound(ja + rs, ja + re, c - 1) - ja);
                on_slot(slot, vidx(e, o + k * nb + j));
              }
              else if(c < 0)
                on_rhs(vidx(e, o + k * nb + j), F->h_fixval[t.dof][-c - 1]);
            }
          else
            on_rhs(vidx(e, o + k), 1.0);
        }
        o += bil ? nb * nb : nb;
      }
    }
  };
  const long nnz = F->nnz;
  std::vector<int> cptr((size_t)nnz + 1, 0), rptr((size_t)N + 1, 0);
  parallel_for(N, [&](int lo, int hi) {
    for(int r = lo; r < hi; r++)
      visit(r, [&](int slot, int) { cptr[(size_t)slot + 1]++; }, [&](int, double) { rptr[(size_t)r + 1]++; });
  });
  for(long i = 0; i < nnz; i++) cptr[i + 1] += cptr[i];
  for(int i = 0; i < N; i++) rptr[i + 1] += rptr[i];
  std::vector<int> cidx((size_t)std::max(1, cptr[nnz])), ridx((size_t)std::max(1, rptr[N]));
  std::vector<double> rfv((size_t)std::max(1, rptr[N]));
  parallel_for(N, [&](int lo, int hi) {
    std::vector<int> cur;
    for(int r = lo; r < hi; r++) {
      cur.assign(cptr.begin() + ia[r], cptr.begin() + ia[r + 1]);
      int rc = rptr[r];
      visit(r, [&](int slot, int idx) { cidx[(size_t)cur[slot - ia[r]]++] = idx; },
            [&](int idx, double fv) {
              ridx[(size_t)rc] = idx;
              rfv[(size_t)rc++] = fv;
            });
    }
  });
  GDP_TRY(dalloc(&F->d_ek_cb, cb.size()));
  GDP_TRY(h2d(F->d_ek_cb, cb.data(), cb.size()));
  GDP_TRY(dalloc(&F->d_ek, (size_t)std::max(1L, cb[nch])));
  GDP_TRY(dalloc(&F->d_cptr, cptr.size()));
  GDP_TRY(h2d(F->d_cptr, cptr.data(), cptr.size()));
  GDP_TRY(dalloc(&F->d_cidx, cidx.size()));
  GDP_TRY(h2d(F->d_cidx, cidx.data(), cidx.size()));
  GDP_TRY(dalloc(&F->d_rptr, rptr.size()));
  GDP_TRY(h2d(F->d_rptr, rptr.data(), rptr.size()));
  GDP_TRY(dalloc(&F->d_ridx, ridx.size()));
  GDP_TRY(h2d(F->d_ridx, ridx.data(), ridx.size()));
  GDP_TRY(dalloc(&F->d_rfv, rfv.size()));
  GDP_TRY(h2d(F->d_rfv, rfv.data(), rfv.size()));
  GDP_TRY(sync());
  return GDP_OK;
}

// Mesh, geometry buffer, fixed values, coefficient and activity tables, and
// the plan's fields that do not depend on the DOF numbering.
static int fe_upload_mesh(gdp_fe *F)
{
  const int E = F->nelem, nspaces = F->nspaces, nterms = F->nterms;
  GDP_TRY(dalloc(&F->d_etype, E));
  GDP_TRY(dalloc(&F->d_region, E));
  GDP_TRY(dalloc(&F->d_conn, (size_t)E * 4));
  GDP_TRY(dalloc(&F->d_geokind, E));
  GDP_TRY(dalloc(&F->d_xyz, (size_t)F->nnodes * 3));
  GDP_TRY(dalloc(&F->d_geo, (size_t)E * kGeoStride));
  GDP_TRY(h2d(F->d_etype, F->h_etype.data(), E));
  GDP_TRY(h2d(F->d_region, F->h_region.data(), E));
  GDP_TRY(h2d(F->d_conn, F->h_conn.data(), (size_t)E * 4));
  GDP_TRY(h2d(F->d_geokind, F->h_geokind.data(), E));
  GDP_TRY(h2d(F->d_xyz, F->h_xyz.data(), (size_t)F->nnodes * 3));
  if(F->family == FAM_EDGE) {
    GDP_TRY(dalloc(&F->d_edges, (size_t)E * 6));
    GDP_TRY(h2d(F->d_edges, F->h_edges.data(), (size_t)E * 6));
  }
  for(int s = 0; s < nspaces; s++) {
    GDP_TRY(dalloc(&F->d_fixval[s], F->h_fixval[s].size()));
    GDP_TRY(h2d(F->d_fixval[s], F->h_fixval[s].data(), F->h_fixval[s].size()));
  }
  DevPlan &P = F->plan;
  P.family = F->family;
  P.nspaces = nspaces;
  P.nterms = nterms;
  P.nreg = F->nreg;
  P.nmax = F->nmax;
  P.etype = F->d_etype;
  P.region = F->d_region;
  P.conn = F->d_conn;
  P.edges = F->d_edges;
  P.geokind = F->d_geokind;
  P.geo = F->d_geo;
  for(int s = 0; s < nspaces; s++) {
    P.fixval[s] = F->d_fixval[s];
    P.bf[s] = F->h_bf[s];
  }
  for(int ti = 0; ti < nterms; ti++) {
    const gdp_term_desc &t = F->h_terms[ti];
    GDP_TRY(dalloc(&F->d_coef[ti], F->h_coef[ti].size()));
    GDP_TRY(h2d(F->d_coef[ti], F->h_coef[ti].data(), F->h_coef[ti].size()));
    GDP_TRY(dalloc(&F->d_active[ti], (size_t)F->nreg));
    GDP_TRY(h2d(F->d_active[ti], F->h_active[ti].data(), (size_t)F->nreg));
    DevTerm &d = P.term[ti];
    d.kind = t.kind;
    d.equ = t.equ;
    d.dof = t.dof;
    d.npoints = t.npoints;
    d.ncoef = t.ncoef;
    d.perp = F->h_perp[ti];
    d.coef = F->d_coef[ti];
    d.active = F->d_active[ti];
  }
  P.nelem = E;
  GDP_TRY(sync());
  return GDP_OK;
}

// The plan's pattern fields (numbering, incidence, CSR, two-phase tables).
static void fe_plan_pattern(gdp_fe *F)
{
  DevPlan &P = F->plan;
  P.N = F->N;
  for(int s = 0; s < F->nspaces; s++) P.dofnum[s] = F->d_dofnum[s];
  P.inc_ptr = F->d_inc_ptr;
  P.inc_item = F->d_inc_item;
  P.row_space = F->d_row_space;
  P.row_ent = F->d_row_ent;
  P.ia = F->d_ia;
  P.ja = F->d_ja;
  P.ek_cb = F->d_ek_cb;
}

// Device copies on the first Generate after a host PRE pass.
static int fe_upload(gdp_fe *F)
{
  if(F->uploaded) return GDP_OK;
  GDP_TRY(ensure_ready());
  const int N = F->N, nmax = F->nmax, nspaces = F->nspaces;
  GDP_TRY(fe_upload_mesh(F));
  for(int s = 0; s < nspaces; s++) {
    GDP_TRY(dalloc(&F->d_dofnum[s], (size_t)nmax + 1));
    GDP_TRY(h2d(F->d_dofnum[s], F->h_dofnum[s].data(), (size_t)nmax + 1));
  }
  GDP_TRY(dalloc(&F->d_inc_ptr, F->h_inc_ptr.size()));
  GDP_TRY(dalloc(&F->d_inc_item, std::max<size_t>(1, F->h_inc_item.size())));
  GDP_TRY(h2d(F->d_inc_ptr, F->h_inc_ptr.data(), F->h_inc_ptr.size()));
  GDP_TRY(h2d(F->d_inc_item, F->h_inc_item.data(), F->h_inc_item.size()));
  GDP_TRY(dalloc(&F->d_row_space, std::max(1, N)));
  GDP_TRY(dalloc(&F->d_row_ent, std::max(1, N)));
  GDP_TRY(h2d(F->d_row_space, F->h_row_space.data(), (size_t)N));
  GDP_TRY(h2d(F->d_row_ent, F->h_row_ent.data(), (size_t)N));
  GDP_TRY(dalloc(&F->d_ia, (size_t)N + 1));
  GDP_TRY(dalloc(&F->d_ja, std::max<size_t>(1, (size_t)F->nnz)));
  GDP_TRY(h2d(F->d_ia, F->h_ia.data(), (size_t)N + 1));
  GDP_TRY(h2d(F->d_ja, F->h_ja.data(), (size_t)F->nnz));
  GDP_TRY(fe_build_two_phase(F));
  fe_plan_pattern(F);
  for(auto &e : F->ev) GDP_HIP(hipEventCreate(&e));
  GDP_TRY(sync());
  F->uploaded = true;
  return GDP_OK;
}

// In-place exclusive scan of n ints on the library stream.
static int scan_excl(int *d, long n)
{
  if(n <= 0) return GDP_OK;
  const long nb = (n + 2047) / 2048;
  int *bs = nullptr;
  GDP_TRY(dalloc(&bs, (size_t)nb));
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(256), 0, ctx().stream, d, n, bs);
  int rc = hipGetLastError() == hipSuccess ? GDP_OK : set_error(GDP_E_HIP, "k_scan_local launch");
  if(!rc && nb > 1) {
    rc = scan_excl(bs, nb);
    if(!rc) hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(256), 0, ctx().stream, d, n, bs);
  }
  if(!rc) rc = sync();
  dfree(bs);
  return rc;
}

static int read_int(const int *d, int *h) { return d2h(h, d, 1); }

static bool fe_prep_on_device()
{
  // GDP_HOST_PREP=1 keeps the PRE pass on the host
  const char *e = std::getenv("GDP_HOST_PREP");
  return ctx().ready && !(e && e[0] && e[0] != '0');
}

// PRE pass on the device: the host pass's tables, built with marks, scans and
// per-entity / per-row threads; the mesh is uploaded at the same time.  The
// host keeps copies of the numbering and of the pattern (API, renumbering).
static int fe_prep_device(gdp_fe *F)
{
  hipStream_t st = ctx().stream;
  const int E = F->nelem, nmax = F->nmax, nsp = F->nspaces;
  GDP_TRY(fe_upload_mesh(F));
  DevPlan &P = F->plan;
  std::vector<int *> tmp;
  auto talloc = [&](int **p, size_t n) {
    int rc = dalloc(p, std::max<size_t>(1, n));
    if(!rc) tmp.push_back(*p);
    return rc;
  };
  auto done = [&](int rc) {
    (void)sync();
    for(int *p : tmp) dfree(p);
    return rc;
  };
  auto zero = [&](void *p, size_t bytes) {
    return hipMemsetAsync(p, 0, bytes, st) == hipSuccess ? GDP_OK : set_error(GDP_E_HIP, "hipMemsetAsync");
  };
  int *err = nullptr;
  GDP_TRY(talloc(&err, 2));
  const int einit[2] = {0, 0x7fffffff};
  int rc = h2d(err, einit, 2);
  // marks, fixed entities
  PrepOut O{};
  for(int s = 0; s < nsp && !rc; s++) {
    rc = dalloc(&F->d_dofnum[s], (size_t)nmax + 1);
    if(!rc) rc = zero(F->d_dofnum[s], sizeof(int) * ((size_t)nmax + 1));
    O.dofnum[s] = F->d_dofnum[s];
    P.dofnum[s] = F->d_dofnum[s];
  }
  if(rc) return done(rc);
  if(E) hipLaunchKernelGGL(k_prep_mark, dim3((E + 255) / 256), dim3(256), 0, st, P, O, err);
  for(int s = 0; s < nsp && !rc; s++) {
    const int nf = (int)F->h_fixent[s].size();
    if(!nf) continue;
    int *fe_ = nullptr;
    rc = talloc(&fe_, nf);
    if(!rc) rc = h2d(fe_, F->h_fixent[s].data(), (size_t)nf);
    if(!rc) hipLaunchKernelGGL(k_prep_fix, dim3((nf + 255) / 256), dim3(256), 0, st, nf, fe_, nmax, F->d_dofnum[s]);
  }
  int h_err[2] = {0, 0};
  if(!rc) rc = d2h(h_err, err, 2);
  if(rc) return done(rc);
  if(h_err[0] & 1) return done(set_error(GDP_E_ARG, "an element entity is out of range"));
  // numbering in NumType order
  std::vector<int> order(nsp);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return F->h_numtype[x] < F->h_numtype[y]; });
  int *flag = nullptr;
  rc = talloc(&flag, (size_t)nmax + 2);
  if(!rc) rc = dalloc(&F->d_row_space, (size_t)nsp * ((size_t)nmax + 1));
  if(!rc) rc = dalloc(&F->d_row_ent, (size_t)nsp * ((size_t)nmax + 1));
  long N = 0;
  const unsigned gent = (unsigned)((nmax + 2 + 255) / 256);
  for(int s : order) {
    if(rc) break;
    hipLaunchKernelGGL(k_prep_flags, dim3(gent), dim3(256), 0, st, nmax, F->d_dofnum[s], flag);
    rc = scan_excl(flag, (long)nmax + 2);
    int cnt = 0;
    if(!rc) rc = read_int(flag + nmax + 1, &cnt);
    if(!rc)
      hipLaunchKernelGGL(k_prep_number, dim3(gent), dim3(256), 0, st, nmax, s, (int)N, flag, F->d_dofnum[s],
                         F->d_row_space, F->d_row_ent);
    N += cnt;
  }
  if(rc) return done(rc);
  F->N = (int)N;
  P.N = (int)N;
  P.row_space = F->d_row_space;
  P.row_ent = F->d_row_ent;
  // incidence
  int *cur = nullptr;
  rc = dalloc(&F->d_inc_ptr, (size_t)nmax + 2);
  if(!rc) rc = zero(F->d_inc_ptr, sizeof(int) * ((size_t)nmax + 2));
  if(!rc && E) hipLaunchKernelGGL(k_prep_inc, dim3((E + 255) / 256), dim3(256), 0, st, P, F->d_inc_ptr, nullptr, nullptr, false);
  if(!rc) rc = scan_excl(F->d_inc_ptr, (long)nmax + 2);
  int ninc = 0;
  if(!rc) rc = read_int(F->d_inc_ptr + nmax + 1, &ninc);
  if(!rc) rc = dalloc(&F->d_inc_item, (size_t)std::max(1, ninc));
  if(!rc) rc = talloc(&cur, (size_t)nmax + 2);
  if(!rc && hipMemcpyAsync(cur, F->d_inc_ptr, sizeof(int) * ((size_t)nmax + 2), hipMemcpyDeviceToDevice, st) != hipSuccess)
    rc = set_error(GDP_E_HIP, "hipMemcpyAsync");
  if(!rc && E) hipLaunchKernelGGL(k_prep_inc, dim3((E + 255) / 256), dim3(256), 0, st, P, nullptr, cur, F->d_inc_item, true);
  if(!rc) hipLaunchKernelGGL(k_prep_inc_sort, dim3(gent), dim3(256), 0, st, nmax, F->d_inc_ptr, F->d_inc_item);
  if(rc) return done(rc);
  P.inc_ptr = F->d_inc_ptr;
  P.inc_item = F->d_inc_item;
  // pattern
  if(N * (long)kPrepRow > 2147483647L) return done(kPrepFallback);
  rc = dalloc(&F->d_ia, (size_t)N + 1);
  if(!rc) rc = zero(F->d_ia, sizeof(int) * ((size_t)N + 1));
  const unsigned grow = (unsigned)((N + 63) / 64);
  if(!rc && N) hipLaunchKernelGGL(k_prep_pattern<false>, dim3(grow), dim3(64), 0, st, P, F->d_ia, nullptr, err);
  if(!rc) rc = d2h(h_err, err, 2);
  if(rc) return done(rc);
  if(h_err[0] & 2) return done(kPrepFallback);  // a row longer than kPrepRow
  if(h_err[1] != 0x7fffffff) return done(set_error(GDP_E_ZERO_ROW, "Zero row in matrix (DOF %d)", h_err[1] + 1));
  rc = scan_excl(F->d_ia, N + 1);
  int nnz = 0;
  if(!rc) rc = read_int(F->d_ia + N, &nnz);
  if(rc) return done(rc);
  F->nnz = nnz;
  P.ia = F->d_ia;
  rc = dalloc(&F->d_ja, (size_t)std::max(1, nnz));
  P.ja = F->d_ja;
  if(!rc && N) hipLaunchKernelGGL(k_prep_pattern<true>, dim3(grow), dim3(64), 0, st, P, nullptr, F->d_ja, err);
  if(rc) return done(rc);
  // two-phase tables
  const int nch = (E + 63) / 64;
  int *cb = nullptr, *ccur = nullptr;
  rc = talloc(&cb, (size_t)nch + 1);
  if(!rc) rc = zero(cb, sizeof(int) * ((size_t)nch + 1));
  if(!rc && nch) hipLaunchKernelGGL(k_prep_chunks, dim3(nch), dim3(64), 0, st, P, cb);
  if(!rc) rc = scan_excl(cb, (long)nch + 1);
  int ektot = 0;
  if(!rc) rc = read_int(cb + nch, &ektot);
  if(!rc && ektot < 0) rc = set_error(GDP_E_UNSUPPORTED, "element blocks exceed int32 indexing");
  if(!rc) rc = dalloc(&F->d_ek_cb, (size_t)nch + 1);
  if(!rc) hipLaunchKernelGGL(k_prep_widen, dim3((nch + 256) / 256), dim3(256), 0, st, nch, cb, F->d_ek_cb);
  if(!rc) rc = dalloc(&F->d_ek, (size_t)std::max(1, ektot));
  F->ek_total = ektot;
  if(!rc) rc = dalloc(&F->d_cptr, (size_t)nnz + 1);
  if(!rc) rc = zero(F->d_cptr, sizeof(int) * ((size_t)nnz + 1));
  if(!rc) rc = dalloc(&F->d_rptr, (size_t)N + 1);
  if(!rc) rc = zero(F->d_rptr, sizeof(int) * ((size_t)N + 1));
  if(!rc && N)
    hipLaunchKernelGGL(k_prep_visit<false>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, P, cb,
                       F->d_cptr, nullptr, nullptr, F->d_rptr, nullptr, nullptr, nullptr);
  if(!rc) rc = scan_excl(F->d_cptr, (long)nnz + 1);
  if(!rc) rc = scan_excl(F->d_rptr, N + 1);
  int ncon = 0, nrhs = 0;
  if(!rc) rc = read_int(F->d_cptr + nnz, &ncon);
  if(!rc) rc = read_int(F->d_rptr + N, &nrhs);
  if(!rc && (ncon < 0 || nrhs < 0)) rc = set_error(GDP_E_UNSUPPORTED, "contribution lists exceed int32 indexing");
  if(!rc) rc = dalloc(&F->d_cidx, (size_t)std::max(1, ncon));
  if(!rc) rc = dalloc(&F->d_ridx, (size_t)std::max(1, nrhs));
  if(!rc) rc = dalloc(&F->d_rfv, (size_t)std::max(1, nrhs));
  if(!rc) rc = talloc(&ccur, (size_t)nnz + 1);
  if(!rc && hipMemcpyAsync(ccur, F->d_cptr, sizeof(int) * ((size_t)nnz + 1), hipMemcpyDeviceToDevice, st) != hipSuccess)
    rc = set_error(GDP_E_HIP, "hipMemcpyAsync");
  if(!rc && N)
    hipLaunchKernelGGL(k_prep_visit<true>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, P, cb, nullptr,
                       ccur, F->d_cidx, nullptr, F->d_rptr, F->d_ridx, F->d_rfv);
  if(!rc && hipGetLastError() != hipSuccess) rc = set_error(GDP_E_HIP, "PRE pass launch");
  // host copies
  F->h_dofnum.assign(nsp, std::vector<int>((size_t)nmax + 1, 0));
  for(int s = 0; s < nsp && !rc; s++) rc = d2h(F->h_dofnum[s].data(), F->d_dofnum[s], (size_t)nmax + 1);
  F->h_ia.assign((size_t)N + 1, 0);
  F->h_ja.assign((size_t)nnz, 0);
  F->h_row_space.assign((size_t)N, 0);
  F->h_row_ent.assign((size_t)N, 0);
  if(!rc) rc = d2h(F->h_ia.data(), F->d_ia, (size_t)N + 1);
  if(!rc && nnz) rc = d2h(F->h_ja.data(), F->d_ja, (size_t)nnz);
  if(!rc && N) rc = d2h(F->h_row_space.data(), F->d_row_space, (size_t)N);
  if(!rc && N) rc = d2h(F->h_row_ent.data(), F->d_row_ent, (size_t)N);
  if(rc) return done(rc);
  fe_plan_pattern(F);
  for(auto &e : F->ev)
    if(hipEventCreate(&e) != hipSuccess) return done(set_error(GDP_E_HIP, "hipEventCreate"));
  F->uploaded = true;
  F->prepped_on_device = true;
  return done(GDP_OK);
}

extern "C" int gdp_fe_destroy(gdp_fe *fe)
{
  if(!fe) return GDP_OK;
  (void)sync();
  delete fe;
  return GDP_OK;
}

extern "C" int gdp_fe_num_dofs(const gdp_fe *fe, int *N, long *nnz)
{
  if(!fe) return set_error(GDP_E_ARG, "null fe");
  if(N) *N = fe->N;
  if(nnz) *nnz = fe->nnz;
  return GDP_OK;
}

extern "C" int gdp_fe_dof_table(const gdp_fe *fe, int space, int *dofnum, int nmax)
{
  if(!fe || space < 0 || space >= fe->nspaces || !dofnum || nmax != fe->nmax)
    return set_error(GDP_E_ARG, "gdp_fe_dof_table: bad arguments (nmax %d != %d)", nmax,
                     fe ? fe->nmax : -1);
  std::copy(fe->h_dofnum[space].begin(), fe->h_dofnum[space].end(), dofnum);
  return GDP_OK;
}

// The DofData block of a .pre file (Dof_WriteFilePRE, DofData.cpp:363-460):
// header lines, then one line per DOF in (NumType, Entity, Harmonic) order --
// the in-order walk of the DOF tree (Dof_WriteDofPRE, :412-456): unknowns
// "NumType Entity 0 1 NumDof 1", fixed DOFs "NumType Entity 0 2 value 0"
// (value as LinAlg_PrintScalar writes it; constraints without a time function
// store time function index 0, which is also the DofData's only one).
extern "C" int gdp_fe_write_pre(const gdp_fe *fe, FILE *f, int num, int resolution_index, int system_index,
                                int nfs, const int *fs_index)
{
  if(!fe || !f || nfs < 0 || (nfs && !fs_index)) return set_error(GDP_E_ARG, "gdp_fe_write_pre: bad arguments");
  std::vector<int> order(fe->nspaces);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return fe->h_numtype[x] < fe->h_numtype[y]; });
  long nany = 0;
  for(int s = 0; s < fe->nspaces; s++)
    for(int ent = 1; ent <= fe->nmax; ent++) nany += fe->h_dofnum[s][ent] != 0;
  fprintf(f, "$DofData /* #%d */\n", num);
  fprintf(f, "%d %d\n", resolution_index, system_index);
  fprintf(f, "%d", nfs);
  for(int i = 0; i < nfs; i++) fprintf(f, " %d", fs_index[i]);
  fprintf(f, "\n");
  fprintf(f, "1 0\n");  // time function indices
  fprintf(f, "1 0\n");  // harmonics
  fprintf(f, "%ld %d\n", nany, fe->N);
  for(int s : order)
    for(int ent = 1; ent <= fe->nmax; ent++) {
      const int d = fe->h_dofnum[s][ent];
      if(d > 0) fprintf(f, "%d %d 0 1 %d 1\n", fe->h_numtype[s], ent, d);
      else if(d < 0) fprintf(f, "%d %d 0 2 %.16g 0\n", fe->h_numtype[s], ent, fe->h_fixval[s][-d - 1]);
    }
  fprintf(f, "$EndDofData\n");
  fflush(f);
  return ferror(f) ? set_error(GDP_E_ARG, "gdp_fe_write_pre: write error") : GDP_OK;
}

// element pass, then the CSR-entry and RHS sums (enqueue only)
static int fe_two_phase(gdp_fe *fe, double *a, double *b)
{
  hipStream_t s = ctx().stream;
  if(fe->nelem) {
    dim3 ge((fe->nelem + 127) / 128);
    if(fe->family == FAM_NODAL) hipLaunchKernelGGL(k_elem_nodal, ge, dim3(128), 0, s, fe->plan, fe->d_ek);
    else hipLaunchKernelGGL(k_elem_edge, ge, dim3(128), 0, s, fe->plan, fe->d_ek);
  }
  if(fe->nnz)
    hipLaunchKernelGGL(k_slot_sum, dim3((unsigned)((fe->nnz + 255) / 256)), dim3(256), 0, s, fe->nnz,
                       fe->d_cptr, fe->d_cidx, fe->d_ek, a);
  if(fe->N)
    hipLaunchKernelGGL(k_rhs_sum, dim3((fe->N + 255) / 256), dim3(256), 0, s, fe->N, fe->d_rptr, fe->d_ridx,
                       fe->d_rfv, fe->d_ek, b);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_fe_generate(gdp_fe *fe, gdp_matrix *A, gdp_vector *b)
{
  if(!fe || !A || !b) return set_error(GDP_E_ARG, "gdp_fe_generate: null argument");
  if(A->n != fe->N || b->n != fe->N)
    return set_error(GDP_E_ARG, "gdp_fe_generate: system size %d, matrix %d, vector %d", fe->N,
                     A->n, b->n);
  GDP_TRY(fe_upload(fe));
  hipStream_t s = ctx().stream;
  // adopt the formulation's pattern (LinAlg_CreateMatrix + first Generate)
  if(A->csr.nnz != fe->nnz || A->h_ia != fe->h_ia || A->h_ja != fe->h_ja) {
    A->csr.release();
    A->csr.n = fe->N;
    A->csr.nnz = fe->nnz;
    GDP_TRY(dalloc(&A->csr.ia, (size_t)fe->N + 1));
    GDP_TRY(dalloc(&A->csr.ja, std::max<size_t>(1, (size_t)fe->nnz)));
    GDP_TRY(dalloc(&A->csr.a, std::max<size_t>(1, (size_t)fe->nnz)));
    GDP_HIP(hipMemcpyAsync(A->csr.ia, fe->d_ia, sizeof(int) * (fe->N + 1), hipMemcpyDeviceToDevice, s));
    GDP_HIP(hipMemcpyAsync(A->csr.ja, fe->d_ja, sizeof(int) * fe->nnz, hipMemcpyDeviceToDevice, s));
    A->h_ia = fe->h_ia;
    A->h_ja = fe->h_ja;
    static uint64_t pv = 1ull << 40;
    A->pattern_version = ++pv;
  }
  A->staged.clear();
  b->staged.clear();
  GDP_HIP(hipEventRecord(fe->ev[0], s));
  if(fe->nelem)
    hipLaunchKernelGGL(k_geometry, dim3((fe->nelem + 255) / 256), dim3(256), 0, s, fe->nelem,
                       fe->d_geokind, fe->d_conn, fe->d_xyz, fe->d_geo);
  GDP_HIP(hipGetLastError());
  GDP_HIP(hipEventRecord(fe->ev[1], s));
  GDP_TRY(fe_two_phase(fe, A->csr.a, b->d));
  GDP_HIP(hipEventRecord(fe->ev[2], s));
  GDP_TRY(sync());
  float t01 = 0, t12 = 0;
  GDP_HIP(hipEventElapsedTime(&t01, fe->ev[0], fe->ev[1]));
  GDP_HIP(hipEventElapsedTime(&t12, fe->ev[1], fe->ev[2]));
  fe->ms_geo = t01;
  fe->ms_gather = t12;
  fe->ms_total = t01 + t12;
  static uint64_t vv = 1ull << 50;
  A->value_version = ++vv;
  return GDP_OK;
}

namespace gdp {

// Generate[A] enqueued only: same kernels as gdp_fe_generate, the values
// cleared by a kernel instead of a memset, no events and no host
// synchronisation, so that a caller may capture it in a graph.  A must
// already carry the formulation's pattern (one gdp_fe_generate before).
int fe_generate_enqueue(gdp_fe *fe, gdp_matrix *A, gdp_vector *b)
{
  if(!fe || !A || !b) return set_error(GDP_E_ARG, "fe_generate_enqueue: null argument");
  if(!fe->uploaded || A->csr.nnz != fe->nnz || A->h_ia != fe->h_ia || b->n != fe->N)
    return set_error(GDP_E_STATE, "fe_generate_enqueue: generate once before capturing");
  hipStream_t s = ctx().stream;
  if(fe->nelem)
    hipLaunchKernelGGL(k_geometry, dim3((fe->nelem + 255) / 256), dim3(256), 0, s, fe->nelem,
                       fe->d_geokind, fe->d_conn, fe->d_xyz, fe->d_geo);
  GDP_TRY(fe_two_phase(fe, A->csr.a, b->d));
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}
}  // namespace gdp

extern "C" int gdp_fe_last_timing(const gdp_fe *fe, double *ms_total, double *ms_geometry,
                                  double *ms_gather)
{
  if(!fe) return set_error(GDP_E_ARG, "null fe");
  if(ms_total) *ms_total = fe->ms_total;
  if(ms_geometry) *ms_geometry = fe->ms_geo;
  if(ms_gather) *ms_gather = fe->ms_gather;
  return GDP_OK;
}

extern "C" int gdp_fe_pattern(const gdp_fe *fe, int *ia, int *ja)
{
  if(!fe) return set_error(GDP_E_ARG, "null fe");
  if(ia) std::copy(fe->h_ia.begin(), fe->h_ia.end(), ia);
  if(ja) std::copy(fe->h_ja.begin(), fe->h_ja.end(), ja);
  return GDP_OK;
}

// =============================================================================
// Post-processing: local and integral quantities per element
// (Pos_LocalOrIntegralQuantity, Cal_PostQuantity.cpp:29-247, with the
// interpolation of Pos_FemInterpolation, Pos_FemInterpolation.cpp:226-330)
// =============================================================================
namespace {

struct PostArgs {
  int kind, bf, nsel, ncoef, npoints;
  const int *sel;  // selected elements, ascending
  const int *etype, *region, *conn, *edges;
  const double *xyz;
  const int *dofnum;
  const double *fixval, *x;
  const unsigned char *insupport;  // per region: in the space's support
  const double *coef;              // coef[region * ncoef + c]
  double *out;
};

// reference coordinates of vertex i (tri: (0,0) (1,0) (0,1); tet adds (0,0,1))
__device__ inline void vertex_uvw(int i, double &u, double &v, double &w)
{
  u = i == 1 ? 1. : 0.;
  v = i == 2 ? 1. : 0.;
  w = i == 3 ? 1. : 0.;
}

// One element: {u} at each vertex, {d u} (gradient, curl or 2D perpendicular
// curl), coef * {d u}, or Integral of coef * SquNorm[{d u}].  The sum over the
// element's basis functions runs in their order with Val += BF * DofValue
// (Pos_FemInterpolation.cpp:253-308), the DOF value being the unknown's or the
// fixed value (Dof_GetRealDofValue); outside the space's support the element
// has no basis functions and the value stays 0.
__device__ void post_element(const PostArgs &a, int s, double o[12], double &Tout)
{
  const int e = a.sel[s];
  const int et = a.etype[e];
  const int tri = et != GDP_TET;
  const int nn = tri ? 3 : 4;
  double G[20];
  geo_of(tri ? GEO_TRI2D : GEO_TET, a.conn + (long)e * 4, a.xyz, G);
  const int reg = a.region[e];
  const bool edge = a.bf == GDP_BF_EDGE, perp = a.bf == GDP_BF_PERP_EDGE;
  const int nb = !a.insupport[reg] ? 0 : (edge ? 6 : nn);
  double dv[6];
  bool has[6];
  for(int j = 0; j < nb; j++) {
    const int ent = edge ? abs(a.edges[(long)e * 6 + j]) : a.conn[(long)e * 4 + j];
    const int c = a.dofnum[ent];
    has[j] = c != 0;
    dv[j] = c > 0 ? a.x[c - 1] : (c < 0 ? a.fixval[-c - 1] : 0.);
  }
  const double *I = G + 10;
  const double d = G[9];
  // {d u}: constant on a P1 element
  double D[3] = {0., 0., 0.};
  if(a.kind != GDP_PQ_VALUE) {
    for(int j = 0; j < nb; j++) {
      if(!has[j]) continue;
      double r[3], b[3];
      if(edge) {
        bf_curledge(j, a.edges[(long)e * 6 + j] < 0, r);
        form2(G, d, r, b);
      }
      else if(perp) {
        double g[3];
        bf_gradnode(1, j, g);
        b[0] = (g[1] * G[0] + -g[0] * G[3]) / d;
        b[1] = (g[1] * G[1] + -g[0] * G[4]) / d;
        b[2] = 0.;
      }
      else {
        bf_gradnode(tri, j, r);
        form1(I, r, b);
      }
      D[0] += b[0] * dv[j];
      D[1] += b[1] * dv[j];
      D[2] += b[2] * dv[j];
    }
  }
  const double *cf = a.coef ? a.coef + (long)reg * a.ncoef : nullptr;
  if(a.kind == GDP_PQ_INT_COEF_D2) {
    const double adet = fabs(d);
    const double tmp = cf[0] * (D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    double T = 0.;
    for(int q = 0; q < a.npoints; q++) {
      double u, v, w, wt;
      gauss_pt(tri, a.npoints, q, u, v, w, wt);
      T = T + tmp * (wt * adet);
    }
    Tout = T;
    return;
  }
  if(a.kind == GDP_PQ_D || a.kind == GDP_PQ_COEF_D) {
    double R[3] = {D[0], D[1], D[2]};
    if(a.kind == GDP_PQ_COEF_D) coef_times(a.ncoef, cf, D, R);
    for(int i = 0; i < nn; i++)
      for(int c = 0; c < 3; c++) o[i * 3 + c] = R[c];
    return;
  }
  for(int i = 0; i < nn; i++) {  // GDP_PQ_VALUE at vertex i
    double u, v, w;
    vertex_uvw(i, u, v, w);
    double V[3] = {0., 0., 0.};
    for(int j = 0; j < nb; j++) {
      if(!has[j]) continue;
      if(edge) {
        double r[3], b[3];
        bf_edge(j, u, v, w, a.edges[(long)e * 6 + j] < 0, r);
        form1(I, r, b);
        V[0] += b[0] * dv[j];
        V[1] += b[1] * dv[j];
        V[2] += b[2] * dv[j];
      }
      else if(perp) {  // (0, 0, BF_Node) through ChangeOfCoord_Form1P
        const double b2 = bf_node(1, j, u, v, w) / G[8];
        V[0] += 0. * dv[j];
        V[1] += 0. * dv[j];
        V[2] += b2 * dv[j];
      }
      else
        V[0] += bf_node(tri, j, u, v, w) * dv[j];
    }
    o[i * 3 + 0] = V[0];
    o[i * 3 + 1] = V[1];
    o[i * 3 + 2] = V[2];
  }
}

// One thread per selected element; the 12 values per element leave through
// LDS so that the block's stores are contiguous.
__global__ void __launch_bounds__(128) k_post(PostArgs a)
{
  __shared__ double st[128 * 12];
  const int s = blockIdx.x * 128 + threadIdx.x;
  double o[12];
  for(int i = 0; i < 12; i++) o[i] = 0.;
  double T = 0.;
  if(s < a.nsel) post_element(a, s, o, T);
  if(a.kind == GDP_PQ_INT_COEF_D2) {  // uniform across the grid
    if(s < a.nsel) a.out[s] = T;
    return;
  }
  for(int i = 0; i < 12; i++) st[threadIdx.x * 12 + i] = o[i];
  __syncthreads();
  const long base = (long)blockIdx.x * 128 * 12, lim = (long)a.nsel * 12;
  for(int k = threadIdx.x; k < 128 * 12; k += 128)
    if(base + k < lim) a.out[base + k] = st[k];
}

// Sum of the per-element integrals in two deterministic stages: block b adds
// elements b*256*kPostPer ... in a fixed order (each thread kPostPer
// consecutive values, then the block's tree), one block adds the partials.
constexpr int kPostPer = 16;
__global__ void __launch_bounds__(256) k_post_part(int n, const double *v, double *part)
{
  const long b = ((long)blockIdx.x * 256 + threadIdx.x) * kPostPer;
  double t = 0.;
  for(int i = 0; i < kPostPer; i++)
    if(b + i < n) t += v[b + i];
  t = block_sum<double, 256>(t);
  if(threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256) k_post_total(int np, const double *part, double *total)
{
  const double t = sum_partials<double, 256>(part, np);
  if(threadIdx.x == 0) *total = t;
}

}  // namespace

extern "C" int gdp_fe_post(gdp_fe *fe, const gdp_vector *x, const gdp_post_desc *q, int *nsel,
                           double *out, long out_len, double *total, double *kernel_ms)
{
  if(!fe || !x || !q || !nsel) return set_error(GDP_E_ARG, "gdp_fe_post: null argument");
  if(q->kind < GDP_PQ_VALUE || q->kind > GDP_PQ_INT_COEF_D2)
    return set_error(GDP_E_UNSUPPORTED, "gdp_fe_post: quantity kind %d", q->kind);
  if(q->space < 0 || q->space >= fe->nspaces) return set_error(GDP_E_ARG, "gdp_fe_post: space %d", q->space);
  if(x->n != fe->N) return set_error(GDP_E_ARG, "gdp_fe_post: vector size %d, system %d", x->n, fe->N);
  GDP_TRY(vec_flush(x));
  const bool needc = q->kind == GDP_PQ_COEF_D || q->kind == GDP_PQ_INT_COEF_D2;
  if(needc) {
    if(!q->coef || q->coef_nreg < fe->nreg) return set_error(GDP_E_ARG, "gdp_fe_post: coefficient table too short");
    if(q->kind == GDP_PQ_INT_COEF_D2 ? q->ncoef != 1 : (q->ncoef != 1 && q->ncoef != 9))
      return set_error(GDP_E_UNSUPPORTED, "gdp_fe_post: %d coefficient values", q->ncoef);
  }
  if(q->kind == GDP_PQ_INT_COEF_D2 && q->npoints < 1) return set_error(GDP_E_ARG, "gdp_fe_post: Gauss points");
  const int bf = fe->h_bf[q->space];
  // OnElementsOf / In: the elements of the listed regions, in element order
  std::vector<unsigned char> inreg(fe->nreg, 0);
  for(int r = 0; r < q->nregions; r++)
    if(q->regions && q->regions[r] >= 0 && q->regions[r] < fe->nreg) inreg[q->regions[r]] = 1;
  std::vector<int> sel;
  for(int e = 0; e < fe->nelem; e++) {
    if(!inreg[fe->h_region[e]]) continue;
    const int et = fe->h_etype[e];
    if(et != GDP_TET && et != GDP_TRI) return set_error(GDP_E_UNSUPPORTED, "gdp_fe_post: element type %d", et);
    if(bf == GDP_BF_EDGE && et != GDP_TET) return set_error(GDP_E_UNSUPPORTED, "gdp_fe_post: edges on triangles");
    if(bf == GDP_BF_PERP_EDGE && et != GDP_TRI)
      return set_error(GDP_E_UNSUPPORTED, "gdp_fe_post: perpendicular edges on tetrahedra");
    if(q->kind == GDP_PQ_INT_COEF_D2 &&
       (et == GDP_TET ? (q->npoints != 1 && q->npoints != 4) : (q->npoints != 1 && q->npoints != 3 && q->npoints != 4)))
      return set_error(GDP_E_UNSUPPORTED, "gdp_fe_post: %d Gauss points", q->npoints);
    sel.push_back(e);
  }
  const int ns = (int)sel.size();
  const long per = q->kind == GDP_PQ_INT_COEF_D2 ? 1 : 12;
  *nsel = ns;
  if(out && out_len < per * ns) return set_error(GDP_E_ARG, "gdp_fe_post: output holds %ld values, %ld needed", out_len, per * ns);
  GDP_TRY(fe_upload(fe));
  hipStream_t s = ctx().stream;
  int *d_sel = nullptr;
  double *d_out = nullptr, *d_tot = nullptr, *d_coef = nullptr;
  unsigned char *d_sup = nullptr;
  const int nparts = (int)std::max(1L, ((long)ns + 256L * kPostPer - 1) / (256L * kPostPer));
  int rc = dalloc(&d_sel, (size_t)std::max(1, ns));
  if(!rc) rc = dalloc(&d_out, (size_t)std::max(1L, per * ns) + 1 + nparts);
  if(!rc) rc = dalloc(&d_sup, (size_t)fe->nreg);
  if(!rc && needc) rc = dalloc(&d_coef, (size_t)fe->nreg * q->ncoef);
  if(!rc && ns) rc = h2d(d_sel, sel.data(), (size_t)ns);
  if(!rc) rc = h2d(d_sup, fe->h_support[q->space].data(), (size_t)fe->nreg);
  if(!rc && needc) rc = h2d(d_coef, q->coef, (size_t)fe->nreg * q->ncoef);
  d_tot = d_out + per * ns;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if(!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = set_error(GDP_E_HIP, "hipEventCreate");
  if(!rc) {
    PostArgs a;
    a.kind = q->kind;
    a.bf = bf;
    a.nsel = ns;
    a.ncoef = needc ? q->ncoef : 0;
    a.npoints = q->npoints;
    a.sel = d_sel;
    a.etype = fe->d_etype;
    a.region = fe->d_region;
    a.conn = fe->d_conn;
    a.edges = fe->d_edges;
    a.xyz = fe->d_xyz;
    a.dofnum = fe->d_dofnum[q->space];
    a.fixval = fe->d_fixval[q->space];
    a.x = x->d;
    a.insupport = d_sup;
    a.coef = d_coef;
    a.out = d_out;
    (void)hipEventRecord(e0, s);
    if(ns) hipLaunchKernelGGL(k_post, dim3((ns + 127) / 128), dim3(128), 0, s, a);
    if(q->kind == GDP_PQ_INT_COEF_D2) {
      hipLaunchKernelGGL(k_post_part, dim3(nparts), dim3(256), 0, s, ns, d_out, d_tot + 1);
      hipLaunchKernelGGL(k_post_total, dim3(1), dim3(256), 0, s, nparts, d_tot + 1, d_tot);
    }
    (void)hipEventRecord(e1, s);
    if(hipGetLastError() != hipSuccess) rc = set_error(GDP_E_HIP, "k_post launch");
  }
  if(!rc) rc = sync();
  if(!rc && kernel_ms) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *kernel_ms = ms;
  }
  if(!rc && out && ns) rc = d2h(out, d_out, (size_t)(per * ns));
  if(!rc && total) {
    *total = 0.;
    if(q->kind == GDP_PQ_INT_COEF_D2) rc = d2h(total, d_tot, 1);
  }
  if(e0) (void)hipEventDestroy(e0);
  if(e1) (void)hipEventDestroy(e1);
  dfree(d_sel);
  dfree(d_out);
  dfree(d_sup);
  dfree(d_coef);
  return rc;
}
