// =====================================================================================
//  oracle.cpp -- TEST INFRASTRUCTURE ONLY (the parity checker, never the product).
//
//  An FP64 CPU restatement of FiniteElementContainers.jl's element-assembly path:
//  structured meshes, DOF numbering, sparsity pattern (Is/Js/unknown_dofs/permutation),
//  the SparseArrays.sparse! COO->CSC->CSR reduction, and the element loops for the
//  residual, the element-private stiffness storage and both Jacobian actions.
//  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
//  Pinning of each function (what checks it, besides the product):
//   * structured meshes     : KAT test/TestMesh.jl:99-108 (quad 3x3), node formula checks
//   * DOF numbering         : integer restatement + invariants (tests/test_oracle_pins.py)
//   * pattern + sparse!     : brute force dense assembly on tiny meshes, closed-form nnz
//                             (3n-2)^d NF^2, an independent numpy restatement
//   * shape fns/quadrature  : partition of unity, sum of gradients, volume, patch test
//   * Poisson               : closed form unit-square stiffness, test/poisson/poisson.gold
//   * linear elastic        : closed form + central finite differences of the energy
//   * neo-Hookean           : P(I)=0, P = dpsi/dF and A = dP/dF by central differences
//   * J2 radial return      : yield consistency, elastic predictor, FD of the update,
//                             uniaxial closed form
//   * mass, lumped mass     : closed-form quad4/hex8 element mass matrices, tet10 vertex/edge
//     and diagonals           lumped masses -V/20 and V/5, lumped = row sums of the consistent
//                             mass (test/TestAssemblers.jl:432-521), total mass = rho V,
//                             diagonal = diag of the sparse! CSR, mass action = CSR x v
//   * tet10 node order and the quadrature rule are conventions of this build
//     ("parity unpinned" against the reference's ReferenceFiniteElements.jl).
//
//  Array conventions follow the reference: nodal fields are NF x NN column-major
//  (src/Fields.jl:36-40, 237-254), connectivity is NNPE x NE column-major, 1-based
//  (src/Fields.jl:129-188), quadrature state is NS x NQ x NE (src/Fields.jl:401-455).
// =====================================================================================
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#include <omp.h>

typedef int64_t i64;

namespace {

enum ElemKind { QUAD4 = 0, HEX8 = 1, TET10 = 2 };
enum PhysKind { POISSON = 0, LINEAR_ELASTIC = 1, NEOHOOKEAN = 2, J2 = 3 };

struct RefElem {
  int nnpe, nd, nq;
  double xi[8][3];
  double w[8];
};

// Gauss-Legendre 2-point per direction (quad4, hex8); 4-point degree-2 rule on the
// unit tetrahedron (Keast/Stroud: a=(5-sqrt5)/20, b=(5+3sqrt5)/20, w=1/24).
RefElem make_ref(int kind) {
  RefElem r{};
  const double g = 1.0 / std::sqrt(3.0);
  if (kind == QUAD4) {
    r.nnpe = 4; r.nd = 2; r.nq = 4;
    int q = 0;
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) { r.xi[q][0] = i ? g : -g; r.xi[q][1] = j ? g : -g; r.w[q] = 1.0; ++q; }
  } else if (kind == HEX8) {
    r.nnpe = 8; r.nd = 3; r.nq = 8;
    int q = 0;
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          r.xi[q][0] = i ? g : -g; r.xi[q][1] = j ? g : -g; r.xi[q][2] = k ? g : -g; r.w[q] = 1.0; ++q;
        }
  } else {
    r.nnpe = 10; r.nd = 3; r.nq = 4;
    const double a = (5.0 - std::sqrt(5.0)) / 20.0, b = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0;
    const double pts[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    for (int q = 0; q < 4; ++q) { for (int d = 0; d < 3; ++d) r.xi[q][d] = pts[q][d]; r.w[q] = 1.0 / 24.0; }
  }
  return r;
}

// Lagrange shape functions. N[a], dN[a*nd + j] = dN_a/dxi_j.
// quad4/hex8 nodes follow the structured mesh's local order (StructuredMesh.jl:116-123, 246-249);
// tet10 follows the Exodus TETRA10 order (vertices 1-4, then edges 12,23,31,14,24,34).
void shape(int kind, const double* xi, double* N, double* dN) {
  if (kind == QUAD4) {
    const double s[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    for (int a = 0; a < 4; ++a) {
      double fx = 1 + s[a][0] * xi[0], fy = 1 + s[a][1] * xi[1];
      N[a] = 0.25 * fx * fy;
      dN[a * 2 + 0] = 0.25 * s[a][0] * fy;
      dN[a * 2 + 1] = 0.25 * fx * s[a][1];
    }
  } else if (kind == HEX8) {
    const double s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                            {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    for (int a = 0; a < 8; ++a) {
      double fx = 1 + s[a][0] * xi[0], fy = 1 + s[a][1] * xi[1], fz = 1 + s[a][2] * xi[2];
      N[a] = 0.125 * fx * fy * fz;
      dN[a * 3 + 0] = 0.125 * s[a][0] * fy * fz;
      dN[a * 3 + 1] = 0.125 * fx * s[a][1] * fz;
      dN[a * 3 + 2] = 0.125 * fx * fy * s[a][2];
    }
  } else {
    const double L[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
    const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 4; ++a) {
      N[a] = L[a] * (2 * L[a] - 1);
      for (int j = 0; j < 3; ++j) dN[a * 3 + j] = (4 * L[a] - 1) * dL[a][j];
    }
    const int e[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    for (int m = 0; m < 6; ++m) {
      int p = e[m][0], q = e[m][1];
      N[4 + m] = 4 * L[p] * L[q];
      for (int j = 0; j < 3; ++j) dN[(4 + m) * 3 + j] = 4 * (dL[p][j] * L[q] + L[p] * dL[q][j]);
    }
  }
}

// isoparametric map (ReferenceFiniteElements MappedH1OrL2Interpolants, called at src/Physics.jl:86-92)
// J_ij = sum_a x_a,i dN_a,j ; gradN_X[a*nd+j] = sum_k dN_a,k Jinv_kj ; JxW = det J * w
struct Mapped {
  double N[10], G[10 * 3], Xq[3], JxW;
};

double inv_nd(int nd, const double* J, double* Ji) {
  if (nd == 2) {
    double det = J[0] * J[3] - J[1] * J[2];
    Ji[0] = J[3] / det; Ji[1] = -J[1] / det; Ji[2] = -J[2] / det; Ji[3] = J[0] / det;
    return det;
  }
  double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
  double det = J[0] * c00 + J[1] * c01 + J[2] * c02;
  Ji[0] = c00 / det; Ji[1] = (J[2] * J[7] - J[1] * J[8]) / det; Ji[2] = (J[1] * J[5] - J[2] * J[4]) / det;
  Ji[3] = c01 / det; Ji[4] = (J[0] * J[8] - J[2] * J[6]) / det; Ji[5] = (J[2] * J[3] - J[0] * J[5]) / det;
  Ji[6] = c02 / det; Ji[7] = (J[1] * J[6] - J[0] * J[7]) / det; Ji[8] = (J[0] * J[4] - J[1] * J[3]) / det;
  return det;
}

void map_qp(int kind, const RefElem& re, int q, const double* xel /*nnpe x nd*/, Mapped& m) {
  const int nd = re.nd, nn = re.nnpe;
  double dN[30];
  shape(kind, re.xi[q], m.N, dN);
  double J[9] = {0};
  for (int a = 0; a < nn; ++a)
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j) J[i * nd + j] += xel[a * nd + i] * dN[a * nd + j];
  double Ji[9];
  double det = inv_nd(nd, J, Ji);
  for (int a = 0; a < nn; ++a)
    for (int j = 0; j < nd; ++j) {
      double s = 0;
      for (int k = 0; k < nd; ++k) s += dN[a * nd + k] * Ji[k * nd + j];
      m.G[a * nd + j] = s;
    }
  for (int i = 0; i < nd; ++i) {
    double s = 0;
    for (int a = 0; a < nn; ++a) s += m.N[a] * xel[a * nd + i];
    m.Xq[i] = s;
  }
  m.JxW = det * re.w[q];
}

// ----------------------------------------------------------------- materials
// grad[i*nd+j] = du_i/dX_j ; P[i*nd+j] ; A[(i*nd+j)*(nf*nd) + k*nd+l] = dP_ij/d grad_kl
struct Material {
  int kind;
  double p[8];
};

double source_term(const Material& m, int nd, const double* X) {
  // props: p[0] = source kind (0 constant, 1 = nd*pi^2*prod sin(pi x_d)), p[1] = coefficient
  if ((int)m.p[0] == 0) return m.p[1];
  double s = m.p[1] * nd * M_PI * M_PI;
  for (int d = 0; d < nd; ++d) s *= std::sin(M_PI * X[d]);
  return s;
}

// J2 small-strain plasticity, linear isotropic hardening, one-step radial return
// (Simo & Hughes, Computational Inelasticity, Box 3.1-3.2). State: plastic strain in
// SymmetricTensor order (11,21,31,22,32,33) then eqps (test/mechanics_with_state/TestMechanicsWithState.jl:66-67).
void j2_update(const double* p, const double* grad, const double* st_old, double* st_new, double* sig /*3x3*/,
               double* C /*81, nullable*/) {
  const double E = p[0], nu = p[1], sy = p[2], H = p[3];
  const double mu = E / (2 * (1 + nu)), kap = E / (3 * (1 - 2 * nu));
  double ep[9];
  ep[0] = st_old[0]; ep[1] = ep[3] = st_old[1]; ep[2] = ep[6] = st_old[2];
  ep[4] = st_old[3]; ep[5] = ep[7] = st_old[4]; ep[8] = st_old[5];
  const double alpha = st_old[6];
  double ee[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ee[i * 3 + j] = 0.5 * (grad[i * 3 + j] + grad[j * 3 + i]) - ep[i * 3 + j];
  const double tr = ee[0] + ee[4] + ee[8];
  double s[9];
  for (int i = 0; i < 9; ++i) s[i] = 2 * mu * ee[i];
  for (int i = 0; i < 3; ++i) s[i * 4] -= 2 * mu * tr / 3.0;
  double sn = 0;
  for (int i = 0; i < 9; ++i) sn += s[i] * s[i];
  sn = std::sqrt(sn);
  const double r23 = std::sqrt(2.0 / 3.0);
  const double f = sn - r23 * (sy + H * alpha);
  double dg = 0, nhat[9] = {0};
  if (f > 0) {
    dg = f / (2 * mu + 2.0 * H / 3.0);
    for (int i = 0; i < 9; ++i) nhat[i] = s[i] / sn;
  }
  for (int i = 0; i < 9; ++i) sig[i] = s[i] - 2 * mu * dg * nhat[i];
  for (int i = 0; i < 3; ++i) sig[i * 4] += kap * tr;
  if (st_new) {
    double epn[9];
    for (int i = 0; i < 9; ++i) epn[i] = ep[i] + dg * nhat[i];
    st_new[0] = epn[0]; st_new[1] = epn[3]; st_new[2] = epn[6]; st_new[3] = epn[4]; st_new[4] = epn[7];
    st_new[5] = epn[8]; st_new[6] = alpha + r23 * dg;
  }
  if (C) {
    double th = 1.0, thb = 0.0;
    if (f > 0) {
      th = 1.0 - 2 * mu * dg / sn;
      thb = 1.0 / (1.0 + H / (3 * mu)) - (1.0 - th);
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            double dij = i == j, dkl = k == l, dik = i == k, djl = j == l, dil = i == l, djk = j == k;
            C[(i * 3 + j) * 9 + k * 3 + l] = kap * dij * dkl + 2 * mu * th * (0.5 * (dik * djl + dil * djk) - dij * dkl / 3.0) -
                                             2 * mu * thb * nhat[i * 3 + j] * nhat[k * 3 + l];
          }
  }
}

void material_eval(const Material& m, int nd, int nf, const double* grad, const double* Xq, const double* st_old,
                   double* st_new, double* P, double* src, double* A) {
  const int D = nf * nd;
  if (m.kind == POISSON) {
    for (int j = 0; j < nd; ++j) P[j] = grad[j];
    src[0] = source_term(m, nd, Xq);
    if (A)
      for (int i = 0; i < D * D; ++i) A[i] = (i % (D + 1)) == 0 ? 1.0 : 0.0;
    return;
  }
  for (int i = 0; i < nf; ++i) src[i] = 0.0;
  if (m.kind == LINEAR_ELASTIC && nd == 2) {
    // plane strain (Mechanics(PlaneStrain()), test/mechanics/TestMechanics.jl:24): the 2 x 2 gradient embedded in a
    // 3 x 3 one with a zero third row and column (modify_field_gradients, src/Formulations.jl:421-427), the 3D
    // material evaluated on it, and its stress and tangent restricted to the in-plane components (extract_stress,
    // extract_stiffness, Formulations.jl:391-419)
    double g3[9] = {0}, P3[9], A3[81], s3[3];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) g3[i * 3 + j] = grad[i * 2 + j];
    material_eval(m, 3, 3, g3, Xq, st_old, st_new, P3, s3, A ? A3 : nullptr);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) P[i * 2 + j] = P3[i * 3 + j];
    if (A)
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          for (int k = 0; k < 2; ++k)
            for (int l = 0; l < 2; ++l) A[(i * 2 + j) * 4 + k * 2 + l] = A3[(i * 3 + j) * 9 + k * 3 + l];
    return;
  }
  if (m.kind == LINEAR_ELASTIC) {
    // psi = K/2 tr(eps)^2 + G dev(eps):dev(eps)  (test/mechanics/TestMechanicsCommon.jl:14-20)
    const double K = m.p[1], G = m.p[2];
    double tr = grad[0] + grad[4] + grad[8];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double eps = 0.5 * (grad[i * 3 + j] + grad[j * 3 + i]);
        P[i * 3 + j] = 2 * G * eps + (i == j ? (K - 2.0 * G / 3.0) * tr : 0.0);
      }
    if (A)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l)
              A[(i * 3 + j) * 9 + k * 3 + l] = (K - 2.0 * G / 3.0) * (i == j) * (k == l) + G * ((i == k) * (j == l) + (i == l) * (j == k));
    return;
  }
  if (m.kind == NEOHOOKEAN) {
    // psi = mu/2 (I1 - 3) - mu ln J + lambda/2 (ln J)^2  (compressible neo-Hookean, Bonet & Wood eq. 6.27)
    const double lam = m.p[0], mu = m.p[1];
    double F[9];
    for (int i = 0; i < 9; ++i) F[i] = grad[i] + ((i % 4) == 0 ? 1.0 : 0.0);
    double Fi[9];
    double J = inv_nd(3, F, Fi);
    double lnJ = std::log(J);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) P[i * 3 + j] = mu * (F[i * 3 + j] - Fi[j * 3 + i]) + lam * lnJ * Fi[j * 3 + i];
    if (A)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l)
              A[(i * 3 + j) * 9 + k * 3 + l] = mu * (i == k) * (j == l) + (mu - lam * lnJ) * Fi[l * 3 + i] * Fi[j * 3 + k] +
                                               lam * Fi[j * 3 + i] * Fi[l * 3 + k];
    return;
  }
  // J2: P = sigma (small strain), A = consistent tangent
  double sig[9];
  j2_update(m.p, grad, st_old, st_new, sig, A);
  for (int i = 0; i < 9; ++i) P[i] = sig[i];
}

struct Problem {
  int kind, nd, nf, nnpe, ns;
  i64 nn, ne;
  const double* X;   // nd x nn
  const i64* conn;   // nnpe x ne, 1-based
  Material mat;
  RefElem re;
};

Problem make_problem(int elem_kind, int nd, i64 nn, const double* X, i64 ne, const i64* conn, int nf, int phys,
                     const double* props, int nprops) {
  Problem pr{};
  pr.kind = elem_kind; pr.nd = nd; pr.nf = nf; pr.nn = nn; pr.ne = ne; pr.X = X; pr.conn = conn;
  pr.re = make_ref(elem_kind);
  pr.nnpe = pr.re.nnpe;
  pr.mat.kind = phys;
  for (int i = 0; i < nprops && i < 8; ++i) pr.mat.p[i] = props[i];
  pr.ns = phys == J2 ? 7 : 0;
  return pr;
}

// one element: residual (R_el, nullable), stiffness K_el (NDOFxNDOF row-major, nullable),
// matrix-free action Kv_el (nullable, needs v), new state (nullable)
void element(const Problem& pr, i64 e, const double* U, const double* V, const double* st_old, double* st_new,
             double* Rel, double* Kel, double* KVel) {
  const int nd = pr.nd, nf = pr.nf, nn = pr.nnpe, D = nf * nd, NDOF = nn * nf;
  double xel[30], uel[30], vel[30];
  for (int a = 0; a < nn; ++a) {
    i64 node = pr.conn[e * nn + a] - 1;
    for (int d = 0; d < nd; ++d) xel[a * nd + d] = pr.X[node * nd + d];
    for (int i = 0; i < nf; ++i) {
      uel[a * nf + i] = U[node * nf + i];
      vel[a * nf + i] = V ? V[node * nf + i] : 0.0;
    }
  }
  if (Rel) for (int i = 0; i < NDOF; ++i) Rel[i] = 0;
  if (Kel) for (int i = 0; i < NDOF * NDOF; ++i) Kel[i] = 0;
  if (KVel) for (int i = 0; i < NDOF; ++i) KVel[i] = 0;
  for (int q = 0; q < pr.re.nq; ++q) {
    Mapped m;
    map_qp(pr.kind, pr.re, q, xel, m);
    double grad[9] = {0}, gv[9] = {0};
    for (int a = 0; a < nn; ++a)
      for (int i = 0; i < nf; ++i)
        for (int j = 0; j < nd; ++j) {
          grad[i * nd + j] += uel[a * nf + i] * m.G[a * nd + j];
          gv[i * nd + j] += vel[a * nf + i] * m.G[a * nd + j];
        }
    double P[9], src[3], A[81];
    const double* so = pr.ns ? st_old + (e * pr.re.nq + q) * pr.ns : nullptr;
    double* sn = (pr.ns && st_new) ? st_new + (e * pr.re.nq + q) * pr.ns : nullptr;
    double sn_dummy[7];
    if (pr.ns && !sn) sn = sn_dummy;
    material_eval(pr.mat, nd, nf, grad, m.Xq, so, sn, P, src, (Kel || KVel) ? A : nullptr);
    if (Rel)
      for (int a = 0; a < nn; ++a)
        for (int i = 0; i < nf; ++i) {
          double s = 0;
          for (int j = 0; j < nd; ++j) s += P[i * nd + j] * m.G[a * nd + j];
          Rel[a * nf + i] += m.JxW * (s - m.N[a] * src[i]);
        }
    if (Kel)
      for (int a = 0; a < nn; ++a)
        for (int i = 0; i < nf; ++i)
          for (int b = 0; b < nn; ++b)
            for (int k = 0; k < nf; ++k) {
              double s = 0;
              for (int j = 0; j < nd; ++j)
                for (int l = 0; l < nd; ++l) s += m.G[a * nd + j] * A[(i * nd + j) * D + k * nd + l] * m.G[b * nd + l];
              Kel[(a * nf + i) * NDOF + b * nf + k] += m.JxW * s;
            }
    if (KVel) {
      double dP[9];
      for (int c = 0; c < D; ++c) {
        double s = 0;
        for (int c2 = 0; c2 < D; ++c2) s += A[c * D + c2] * gv[c2];
        dP[c] = s;
      }
      for (int a = 0; a < nn; ++a)
        for (int i = 0; i < nf; ++i) {
          double s = 0;
          for (int j = 0; j < nd; ++j) s += dP[i * nd + j] * m.G[a * nd + j];
          KVel[a * nf + i] += m.JxW * s;
        }
    }
  }
}

// one element's consistent mass M_el[(a,i),(b,k)] = rho sum_q JxW N_a N_b delta_ik (NDOF x NDOF row-major):
// the test physics' mass callbacks (test/mechanics/TestMechanicsCommon.jl:52-85 with rho = props[1],
// test/poisson/TestPoissonCommon.jl:18-41 with rho = 1)
void mass_element(const Problem& pr, i64 e, double rho, double* Mel) {
  const int nd = pr.nd, nf = pr.nf, nn = pr.nnpe, NDOF = nn * nf;
  double xel[30];
  for (int a = 0; a < nn; ++a) {
    i64 node = pr.conn[e * nn + a] - 1;
    for (int d = 0; d < nd; ++d) xel[a * nd + d] = pr.X[node * nd + d];
  }
  for (int i = 0; i < NDOF * NDOF; ++i) Mel[i] = 0;
  for (int q = 0; q < pr.re.nq; ++q) {
    Mapped m;
    map_qp(pr.kind, pr.re, q, xel, m);
    for (int a = 0; a < nn; ++a)
      for (int b = 0; b < nn; ++b)
        for (int i = 0; i < nf; ++i) Mel[(a * nf + i) * NDOF + b * nf + i] += rho * m.JxW * m.N[a] * m.N[b];
  }
}

// one element's row-sum lumped mass m_el[(a,i)] = rho sum_q JxW N_a (TestMechanicsCommon.jl:98-124)
void lumped_element(const Problem& pr, i64 e, double rho, double* mel) {
  const int nd = pr.nd, nf = pr.nf, nn = pr.nnpe;
  double xel[30];
  for (int a = 0; a < nn; ++a) {
    i64 node = pr.conn[e * nn + a] - 1;
    for (int d = 0; d < nd; ++d) xel[a * nd + d] = pr.X[node * nd + d];
  }
  for (int i = 0; i < nn * nf; ++i) mel[i] = 0;
  for (int q = 0; q < pr.re.nq; ++q) {
    Mapped m;
    map_qp(pr.kind, pr.re, q, xel, m);
    for (int a = 0; a < nn; ++a)
      for (int i = 0; i < nf; ++i) mel[a * nf + i] += rho * m.JxW * m.N[a];
  }
}

inline void atomic_add(double* p, double v) {
#pragma omp atomic
  *p += v;
}

}  // namespace

extern "C" {

int orc_num_threads() { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }

// ------------------------------------------------------------------ structured meshes
// StructuredMesh("quad"|"hex", mins, maxs, counts): counts are NODE counts.
// coords: nd x NN col-major; conn: nnpe x NE col-major, 1-based.
// src/meshes/StructuredMesh.jl:85-131 (hex8), 232-255 (quad4), 433-484 (coords, LinRange)
static double linrange(double a, double b, i64 n, i64 i /*1-based*/) {
  // Base.LinRange getindex -> lerpi(j, d, a, b) = (1 - j/d)*a + (j/d)*b
  double t = (double)(i - 1) / (double)(n - 1);
  return (1 - t) * a + t * b;
}

int orc_structured_mesh(int kind, const double* mins, const double* maxs, const i64* counts, double* coords, i64* conn) {
  if (kind == QUAD4) {
    i64 Nx = counts[0], Ny = counts[1];
    i64 n = 0;
    // _two_dimensional_coords!: outer loop over counts[1] (ny), inner over counts[2] (nx) (StructuredMesh.jl:476-477)
    for (i64 ny = 1; ny <= counts[0]; ++ny)
      for (i64 nx = 1; nx <= counts[1]; ++nx) {
        coords[2 * n + 0] = linrange(mins[0], maxs[0], counts[0], nx);
        coords[2 * n + 1] = linrange(mins[1], maxs[1], counts[1], ny);
        ++n;
      }
    auto node = [&](i64 i, i64 j) { return i + (j - 1) * Nx; };
    i64 e = 0;
    for (i64 ex = 1; ex <= Nx - 1; ++ex)
      for (i64 ey = 1; ey <= Ny - 1; ++ey) {
        i64* c = conn + 4 * e;
        c[0] = node(ex, ey); c[1] = node(ex + 1, ey); c[2] = node(ex + 1, ey + 1); c[3] = node(ex, ey + 1);
        ++e;
      }
    return 0;
  }
  if (kind == HEX8) {
    i64 Nx = counts[0], Ny = counts[1], Nz = counts[2];
    i64 n = 0;
    // _three_dimensional_coords!: loops nz over counts[1], ny over counts[2], nx over counts[3] (StructuredMesh.jl:438-440)
    for (i64 nz = 1; nz <= counts[0]; ++nz)
      for (i64 ny = 1; ny <= counts[1]; ++ny)
        for (i64 nx = 1; nx <= counts[2]; ++nx) {
          coords[3 * n + 0] = linrange(mins[0], maxs[0], counts[0], nx);
          coords[3 * n + 1] = linrange(mins[1], maxs[1], counts[1], ny);
          coords[3 * n + 2] = linrange(mins[2], maxs[2], counts[2], nz);
          ++n;
        }
    auto node = [&](i64 i, i64 j, i64 k) { return i + Nx * (j - 1) + Nx * Ny * (k - 1); };
    i64 e = 0;
    for (i64 ex = 1; ex <= Nx - 1; ++ex)
      for (i64 ey = 1; ey <= Ny - 1; ++ey)
        for (i64 ez = 1; ez <= Nz - 1; ++ez) {
          i64* c = conn + 8 * e;
          c[0] = node(ex, ey, ez); c[1] = node(ex + 1, ey, ez); c[2] = node(ex + 1, ey + 1, ez); c[3] = node(ex, ey + 1, ez);
          c[4] = node(ex, ey, ez + 1); c[5] = node(ex + 1, ey, ez + 1); c[6] = node(ex + 1, ey + 1, ez + 1);
          c[7] = node(ex, ey + 1, ez + 1);
          ++e;
        }
    return 0;
  }
  return -1;
}

// ------------------------------------------------------------------ DOF numbering
// update_dofs! without periodic pairs (src/DofManagers.jl:227-298):
// unknown_dofs = 1:N minus sort(dirichlet); dof_to_unknown[d] = rank or -1.
// dirichlet must be the reference's unique(sort(...)) list (src/bcs/DirichletBCs.jl:396-398).
i64 orc_update_dofs(i64 ndofs, const i64* dirichlet, i64 ndir, i64* unknown_dofs, i64* dof_to_unknown) {
  std::vector<char> is_bc(ndofs + 1, 0);
  for (i64 i = 0; i < ndir; ++i) is_bc[dirichlet[i]] = 1;
  i64 k = 0;
  for (i64 d = 1; d <= ndofs; ++d) {
    if (is_bc[d]) { dof_to_unknown[d - 1] = -1; continue; }
    unknown_dofs[k] = d;
    dof_to_unknown[d - 1] = ++k;
  }
  return k;
}

// ------------------------------------------------------------------ sparsity pattern
// Number of reduced COO entries (SparsityPatterns.jl:160-182).
i64 orc_pattern_count(int nf, int nnpe, i64 ne, const i64* conn, const i64* dof_to_unknown) {
  i64 n = 0;
  const int NDOF = nf * nnpe;
  for (i64 e = 0; e < ne; ++e) {
    int nfree = 0;
    for (int a = 0; a < nnpe; ++a)
      for (int d = 0; d < nf; ++d) nfree += dof_to_unknown[nf * (conn[e * nnpe + a] - 1) + d] > 0;
    n += (i64)nfree * nfree;
    (void)NDOF;
  }
  return n;
}

// Reduced SparseMatrixPattern (SparsityPatterns.jl:193-229): Is/Js in unknown numbering,
// unknown_dofs[n] = 1-based slot into the full element storage, permutation = stable
// sortperm of (I<<32)|J (1-based).
void orc_pattern(int nf, int nnpe, i64 ne, const i64* conn, const i64* dof_to_unknown, i64* Is, i64* Js, i64* slots,
                 i64* perm) {
  const int NDOF = nf * nnpe;
  i64 dof_num = 1, n = 0;
  std::vector<i64> dc(NDOF);
  for (i64 e = 0; e < ne; ++e) {
    for (int a = 0; a < nnpe; ++a)
      for (int d = 0; d < nf; ++d) dc[a * nf + d] = nf * (conn[e * nnpe + a] - 1) + d + 1;
    for (int i = 0; i < NDOF; ++i) {
      i64 ri = dof_to_unknown[dc[i] - 1];
      for (int j = 0; j < NDOF; ++j) {
        i64 rj = dof_to_unknown[dc[j] - 1];
        if (ri > 0 && rj > 0) { Is[n] = ri; Js[n] = rj; slots[n] = dof_num; ++n; }
        ++dof_num;
      }
    }
  }
  std::vector<i64> p(n);
  std::iota(p.begin(), p.end(), 0);
  std::stable_sort(p.begin(), p.end(), [&](i64 x, i64 y) {
    i64 kx = (Is[x] << 32) | Js[x], ky = (Is[y] << 32) | Js[y];
    return kx < ky;
  });
  for (i64 i = 0; i < n; ++i) perm[i] = p[i] + 1;
}

// SparseArrays.sparse!(I, J, V, m, n, +, ...) restated (SparsityPatterns.jl:301-308):
// pass 1 counting-sorts entries by row into CSR work arrays and combines duplicates within a
// row at their first occurrence (klasttouch), pass 2 transposes the row-compressed form to CSC.
// Then SparseMatrixCSR(csc) (SparsityPatterns.jl:326-329) gives sorted rows/cols.
// Outputs: rowptr (n+1, 1-based), colval (nnz, 1-based), vals (nnz) and, if cmap != NULL,
// the CSR position (0-based) of each COO entry n -> cmap[n].
// Returns nnz. Call with rowptr/colval/vals NULL to only count.
i64 orc_sparse_csr(i64 nrows, i64 ncoo, const i64* Is, const i64* Js, const double* V, i64* rowptr, i64* colval,
                   double* vals, i64* cmap) {
  const i64 m = nrows;
  // pass 1: counting sort by row
  std::vector<i64> crp(m + 2, 0);
  for (i64 k = 0; k < ncoo; ++k) crp[Is[k] + 1]++;
  crp[1] = 1;
  for (i64 i = 1; i <= m; ++i) crp[i + 1] += crp[i];  // crp[i] = start of row i (1-based)
  std::vector<i64> ccol(ncoo), corig(ncoo);
  std::vector<double> cval(ncoo);
  std::vector<i64> fill(crp.begin(), crp.end());
  for (i64 k = 0; k < ncoo; ++k) {
    i64 r = Is[k];
    i64 pos = fill[r]++;
    ccol[pos - 1] = Js[k];
    cval[pos - 1] = V ? V[k] : 0.0;
    corig[pos - 1] = k;
  }
  // combine duplicates within each row (klasttouch)
  std::vector<i64> klast(m + 1, 0);
  std::vector<i64> where(ncoo);
  i64 wr = 0;
  std::vector<i64> rstart(m + 2);
  for (i64 i = 1; i <= m; ++i) {
    rstart[i] = wr + 1;
    for (i64 p = crp[i]; p < crp[i + 1]; ++p) {
      i64 j = ccol[p - 1];
      if (klast[j] < rstart[i]) {
        ++wr;
        klast[j] = wr;
        ccol[wr - 1] = j;
        cval[wr - 1] = cval[p - 1];
        where[corig[p - 1]] = wr - 1;
      } else {
        cval[klast[j] - 1] += cval[p - 1];
        where[corig[p - 1]] = klast[j] - 1;
      }
    }
  }
  rstart[m + 1] = wr + 1;
  const i64 nnz = wr;
  if (!rowptr) return nnz;
  // pass 2: transpose to CSC (rows sorted within columns), then back to CSR (columns sorted)
  std::vector<i64> ccp(m + 2, 0);
  for (i64 k = 0; k < nnz; ++k) ccp[ccol[k] + 1]++;
  ccp[1] = 1;
  for (i64 j = 1; j <= m; ++j) ccp[j + 1] += ccp[j];
  std::vector<i64> crow(nnz), cfrom(nnz);
  std::vector<double> cv(nnz);
  std::vector<i64> f2(ccp.begin(), ccp.end());
  for (i64 i = 1; i <= m; ++i)
    for (i64 p = rstart[i]; p < rstart[i + 1]; ++p) {
      i64 j = ccol[p - 1];
      i64 pos = f2[j]++;
      crow[pos - 1] = i;
      cv[pos - 1] = cval[p - 1];
      cfrom[pos - 1] = p - 1;
    }
  // CSC -> CSR
  std::vector<i64> rp(m + 2, 0);
  for (i64 k = 0; k < nnz; ++k) rp[crow[k] + 1]++;
  rp[1] = 1;
  for (i64 i = 1; i <= m; ++i) rp[i + 1] += rp[i];
  std::vector<i64> f3(rp.begin(), rp.end());
  std::vector<i64> rowpos_of_rowcomp(nnz);
  for (i64 j = 1; j <= m; ++j)
    for (i64 p = ccp[j]; p < ccp[j + 1]; ++p) {
      i64 i = crow[p - 1];
      i64 pos = f3[i]++;
      colval[pos - 1] = j;
      vals[pos - 1] = cv[p - 1];
      rowpos_of_rowcomp[cfrom[p - 1]] = pos - 1;
    }
  for (i64 i = 0; i <= m; ++i) rowptr[i] = rp[i + 1];
  if (cmap)
    for (i64 k = 0; k < ncoo; ++k) cmap[k] = rowpos_of_rowcomp[where[k]];
  return nnz;
}

// sparse!(I, J, V, n, n, +) alone: the CSC matrix of the default sparse_matrix_type = :csc (_csc_matrix,
// SparsityPatterns.jl:301-308), pass 1 and pass 2 of orc_sparse_csr without the final CSC -> CSR step. Outputs colptr
// (n+1, 1-based), rowval (nnz, 1-based, ascending within each column) and nzval. Returns nnz; call with colptr NULL to
// only count.
i64 orc_sparse_csc(i64 nrows, i64 ncoo, const i64* Is, const i64* Js, const double* V, i64* colptr, i64* rowval,
                   double* nzval) {
  const i64 m = nrows;
  std::vector<i64> crp(m + 2, 0);
  for (i64 k = 0; k < ncoo; ++k) crp[Is[k] + 1]++;
  crp[1] = 1;
  for (i64 i = 1; i <= m; ++i) crp[i + 1] += crp[i];
  std::vector<i64> ccol(ncoo);
  std::vector<double> cval(ncoo);
  std::vector<i64> fill(crp.begin(), crp.end());
  for (i64 k = 0; k < ncoo; ++k) {
    const i64 pos = fill[Is[k]]++;
    ccol[pos - 1] = Js[k];
    cval[pos - 1] = V ? V[k] : 0.0;
  }
  std::vector<i64> klast(m + 1, 0), rstart(m + 2);
  i64 wr = 0;
  for (i64 i = 1; i <= m; ++i) {
    rstart[i] = wr + 1;
    for (i64 p = crp[i]; p < crp[i + 1]; ++p) {
      const i64 j = ccol[p - 1];
      if (klast[j] < rstart[i]) {
        ++wr;
        klast[j] = wr;
        ccol[wr - 1] = j;
        cval[wr - 1] = cval[p - 1];
      } else {
        cval[klast[j] - 1] += cval[p - 1];
      }
    }
  }
  rstart[m + 1] = wr + 1;
  const i64 nnz = wr;
  if (!colptr) return nnz;
  std::vector<i64> ccp(m + 2, 0);
  for (i64 k = 0; k < nnz; ++k) ccp[ccol[k] + 1]++;
  ccp[1] = 1;
  for (i64 j = 1; j <= m; ++j) ccp[j + 1] += ccp[j];
  std::vector<i64> f2(ccp.begin(), ccp.end());
  for (i64 i = 1; i <= m; ++i)
    for (i64 p = rstart[i]; p < rstart[i + 1]; ++p) {
      const i64 pos = f2[ccol[p - 1]]++;
      rowval[pos - 1] = i;
      nzval[pos - 1] = cval[p - 1];
    }
  for (i64 j = 0; j <= m; ++j) colptr[j] = ccp[j + 1];
  return nnz;
}

// Condensed-mode constraint adjustments (use_condensed = true; src/assemblers/Utils.jl:53-167). c is the
// constraint_storage vector: 1 at Dirichlet dofs, 0 elsewhere (SparseMatrixAssembler.jl:93-94,246-248).
// Matrix, compressed form (ptr, idx, nz) of n x n: for every compressed line k (a CSR row, or a CSC column),
// nz = (1 - c[k]) nz, and its diagonal entry gets + penalty c[k], penalty = penalty_scale tr(A) / n with the trace
// taken before the adjustment (Utils.jl:53-145; _sptrace = tr, :43-51). The CSR form zeroes rows, the CSC form
// columns, exactly as the reference's methods for the two types do.
void orc_adjust_matrix(i64 n, const i64* ptr, const i64* idx, double* nz, const double* c, double penalty_scale) {
  double tr = 0.0;
  for (i64 k = 0; k < n; ++k)
    for (i64 p = ptr[k]; p < ptr[k + 1]; ++p)
      if (idx[p - 1] == k + 1) tr += nz[p - 1];
  const double penalty = penalty_scale * tr / (double)n;
  for (i64 k = 0; k < n; ++k)
    for (i64 p = ptr[k]; p < ptr[k + 1]; ++p) {
      nz[p - 1] = (1.0 - c[k]) * nz[p - 1];
      if (idx[p - 1] == k + 1) nz[p - 1] = nz[p - 1] + penalty * c[k];
    }
}

// b = (1 - c) b (Utils.jl:161-167); Av = (1 - c) Av + c v (Utils.jl:150-159).
void orc_adjust_vector(i64 n, double* b, const double* c) {
  for (i64 k = 0; k < n; ++k) b[k] = (1.0 - c[k]) * b[k];
}
void orc_adjust_action(i64 n, double* Av, const double* c, const double* v) {
  for (i64 k = 0; k < n; ++k) Av[k] = (1.0 - c[k]) * Av[k] + c[k] * v[k];
}

// The same CSR as pattern -> gather of the element storage -> orc_sparse_csr, formed row by row so that it runs in
// parallel and needs no COO arrays (full-size checks: 5.1e8 COO entries at hex8 96^3). Row r of unknown dof
// (node n, component d) sums, over n's incident elements in ascending element id and then the element's columns j,
// storage[e NDOF^2 + i NDOF + j] with i = a nf + d (SparsityPatterns.jl:72-84 enumeration): the order in which
// sparse! combines duplicates (its counting sort by row keeps COO order, SparsityPatterns.jl:301-308), so the values
// are bit-identical to orc_sparse_csr's. Columns ascend within each row (SparseMatrixCSR, :326-329).
// Call with colval == NULL to receive rowptr (n_unknowns + 1, 1-based) and return nnz; then again with all arrays.
i64 orc_sparse_csr_rows(int nf, int nnpe, i64 nn, i64 ne, const i64* conn, const i64* dof_to_unknown,
                        i64 n_unknowns, const double* storage, i64* rowptr, i64* colval, double* vals) {
  const int NDOF = nf * nnpe;
  std::vector<i64> noff(nn + 1, 0), nel(ne * nnpe);
  for (i64 k = 0; k < ne * nnpe; ++k) noff[conn[k]]++;
  for (i64 i = 0; i < nn; ++i) noff[i + 1] += noff[i];
  {
    std::vector<i64> fill(noff.begin(), noff.end() - 1);
    for (i64 e = 0; e < ne; ++e)
      for (int a = 0; a < nnpe; ++a) nel[fill[conn[e * nnpe + a] - 1]++] = e * nnpe + a;  // ascending e per node
  }
  const bool count = colval == nullptr;
  if (count) {
    for (i64 r = 0; r <= n_unknowns; ++r) rowptr[r] = 0;
  }
#pragma omp parallel
  {
    std::vector<i64> cols;
    std::vector<i64> mark(nn * nf, -1);  // unknown column -> position in the row
    std::vector<double> acc;
#pragma omp for schedule(dynamic, 256)
    for (i64 n = 0; n < nn; ++n) {
      for (int d = 0; d < nf; ++d) {
        const i64 r = dof_to_unknown[nf * n + d];
        if (r <= 0) continue;
        cols.clear();
        for (i64 t = noff[n]; t < noff[n + 1]; ++t) {
          const i64 e = nel[t] / nnpe;
          for (int b = 0; b < nnpe; ++b)
            for (int k = 0; k < nf; ++k) {
              const i64 c = dof_to_unknown[nf * (conn[e * nnpe + b] - 1) + k];
              if (c > 0 && mark[c - 1] < 0) { mark[c - 1] = 0; cols.push_back(c); }
            }
        }
        std::sort(cols.begin(), cols.end());
        if (count) {
          rowptr[r] = (i64)cols.size();
        } else {
          const i64 base = rowptr[r - 1] - 1;
          acc.assign(cols.size(), 0.0);
          for (size_t q = 0; q < cols.size(); ++q) mark[cols[q] - 1] = (i64)q;
          for (i64 t = noff[n]; t < noff[n + 1]; ++t) {
            const i64 e = nel[t] / nnpe;
            const int i = (int)(nel[t] % nnpe) * nf + d;
            const double* s = storage + e * NDOF * NDOF + (i64)i * NDOF;
            for (int j = 0; j < NDOF; ++j) {
              const i64 c = dof_to_unknown[nf * (conn[e * nnpe + j / nf] - 1) + j % nf];
              if (c > 0) acc[mark[c - 1]] += s[j];
            }
          }
          for (size_t q = 0; q < cols.size(); ++q) {
            colval[base + (i64)q] = cols[q];
            vals[base + (i64)q] = acc[q];
          }
        }
        for (i64 c : cols) mark[c - 1] = -1;
      }
    }
  }
  if (count) {
    rowptr[0] = 1;
    for (i64 r = 1; r <= n_unknowns; ++r) rowptr[r] += rowptr[r - 1];
  }
  return rowptr[n_unknowns] - 1;
}

// ------------------------------------------------------------------ element loops
// Residual into a full NF x NN field (Assemblers.jl:402-432 + :72-87, atomic adds when threaded,
// src/Utils.jl:1-21). U: full field.
void orc_assemble_residual(int elem_kind, int nd, i64 nn, const double* X, i64 ne, const i64* conn, int nf, int phys,
                           const double* props, int nprops, const double* U, const double* st_old, double* st_new,
                           double* R) {
  Problem pr = make_problem(elem_kind, nd, nn, X, ne, conn, nf, phys, props, nprops);
  for (i64 i = 0; i < nn * nf; ++i) R[i] = 0;
  const int NDOF = pr.nnpe * nf;
#pragma omp parallel
  {
    std::vector<double> Rel(NDOF);
#pragma omp for schedule(static)
    for (i64 e = 0; e < ne; ++e) {
      element(pr, e, U, nullptr, st_old, st_new, Rel.data(), nullptr, nullptr);
      for (int d = 0; d < nf; ++d)
        for (int a = 0; a < pr.nnpe; ++a) atomic_add(&R[nf * (conn[e * pr.nnpe + a] - 1) + d], Rel[a * nf + d]);
    }
  }
}

// Element-private stiffness storage (Assemblers.jl:109-124): storage[(e-1)*NDOF^2 + i] = K_el.data[i],
// with K_el.data column-major, i.e. slot (r,c) (r outer, c inner, matching the pattern's i-outer/j-inner
// enumeration SparsityPatterns.jl:76-84) holds K_el[c, r].
void orc_assemble_stiffness_storage(int elem_kind, int nd, i64 nn, const double* X, i64 ne, const i64* conn, int nf,
                                    int phys, const double* props, int nprops, const double* U, const double* st_old,
                                    double* storage) {
  Problem pr = make_problem(elem_kind, nd, nn, X, ne, conn, nf, phys, props, nprops);
  const int NDOF = pr.nnpe * nf;
#pragma omp parallel
  {
    std::vector<double> Kel(NDOF * NDOF);
#pragma omp for schedule(static)
    for (i64 e = 0; e < ne; ++e) {
      element(pr, e, U, nullptr, st_old, nullptr, nullptr, Kel.data(), nullptr);
      double* s = storage + e * NDOF * NDOF;
      for (int r = 0; r < NDOF; ++r)
        for (int c = 0; c < NDOF; ++c) s[r * NDOF + c] = Kel[c * NDOF + r];
    }
  }
}

// Assembled action (MatrixAction.jl:208-238): K_el built, then K_el * v_el, atomic scatter.
// matrix_free != 0: per-qp action (MatrixAction.jl:49-77).
void orc_assemble_action(int elem_kind, int nd, i64 nn, const double* X, i64 ne, const i64* conn, int nf, int phys,
                         const double* props, int nprops, const double* U, const double* V, const double* st_old,
                         int matrix_free, double* KV) {
  Problem pr = make_problem(elem_kind, nd, nn, X, ne, conn, nf, phys, props, nprops);
  const int NDOF = pr.nnpe * nf;
  for (i64 i = 0; i < nn * nf; ++i) KV[i] = 0;
#pragma omp parallel
  {
    std::vector<double> Kel(NDOF * NDOF), out(NDOF), vel(NDOF);
#pragma omp for schedule(static)
    for (i64 e = 0; e < ne; ++e) {
      if (matrix_free) {
        element(pr, e, U, V, st_old, nullptr, nullptr, nullptr, out.data());
      } else {
        element(pr, e, U, nullptr, st_old, nullptr, nullptr, Kel.data(), nullptr);
        for (int a = 0; a < pr.nnpe; ++a)
          for (int i = 0; i < nf; ++i) vel[a * nf + i] = V[nf * (conn[e * pr.nnpe + a] - 1) + i];
        for (int r = 0; r < NDOF; ++r) {
          double s = 0;
          for (int c = 0; c < NDOF; ++c) s += Kel[r * NDOF + c] * vel[c];
          out[r] = s;
        }
      }
      for (int d = 0; d < nf; ++d)
        for (int a = 0; a < pr.nnpe; ++a) atomic_add(&KV[nf * (conn[e * pr.nnpe + a] - 1) + d], out[a * nf + d]);
    }
  }
}

// Consistent mass element storage (assemble_mass!, src/assemblers/Matrix.jl:1-10 -> assemble_matrix!), laid out
// like orc_assemble_stiffness_storage (M_el is symmetric, so its transpose is itself).
void orc_assemble_mass_storage(int elem_kind, int nd, i64 nn, const double* X, i64 ne, const i64* conn, int nf,
                               double rho, double* storage) {
  Problem pr = make_problem(elem_kind, nd, nn, X, ne, conn, nf, POISSON, nullptr, 0);
  const int NDOF = pr.nnpe * nf;
#pragma omp parallel
  {
    std::vector<double> Mel(NDOF * NDOF);
#pragma omp for schedule(static)
    for (i64 e = 0; e < ne; ++e) {
      mass_element(pr, e, rho, Mel.data());
      double* s = storage + e * NDOF * NDOF;
      for (int r = 0; r < NDOF; ++r)
        for (int c = 0; c < NDOF; ++c) s[r * NDOF + c] = Mel[c * NDOF + r];
    }
  }
}

// Row-sum lumped mass into a full NF x NN field (assemble_lumped_mass!, src/assemblers/LumpedMass.jl:29-58).
void orc_assemble_lumped_mass(int elem_kind, int nd, i64 nn, const double* X, i64 ne, const i64* conn, int nf,
                              double rho, double* M) {
  Problem pr = make_problem(elem_kind, nd, nn, X, ne, conn, nf, POISSON, nullptr, 0);
  const int NDOF = pr.nnpe * nf;
  for (i64 i = 0; i < nn * nf; ++i) M[i] = 0;
#pragma omp parallel
  {
    std::vector<double> mel(NDOF);
#pragma omp for schedule(static)
    for (i64 e = 0; e < ne; ++e) {
      lumped_element(pr, e, rho, mel.data());
      for (int d = 0; d < nf; ++d)
        for (int a = 0; a < pr.nnpe; ++a) atomic_add(&M[nf * (conn[e * pr.nnpe + a] - 1) + d], mel[a * nf + d]);
    }
  }
}

// Element diagonals scattered into a full NF x NN field (assemble_diagonal!, src/assemblers/Diagonal.jl:17-68, the
// AssembledDiagonal accumulation Assemblers.jl:42-45): which = 0 the stiffness of (phys, props) at U, which = 1 the
// consistent mass with density rho.
void orc_assemble_diagonal(int elem_kind, int nd, i64 nn, const double* X, i64 ne, const i64* conn, int nf, int phys,
                           const double* props, int nprops, const double* U, const double* st_old, int which,
                           double rho, double* Dg) {
  Problem pr = make_problem(elem_kind, nd, nn, X, ne, conn, nf, phys, props, nprops);
  const int NDOF = pr.nnpe * nf;
  for (i64 i = 0; i < nn * nf; ++i) Dg[i] = 0;
#pragma omp parallel
  {
    std::vector<double> Kel(NDOF * NDOF);
#pragma omp for schedule(static)
    for (i64 e = 0; e < ne; ++e) {
      if (which == 1) mass_element(pr, e, rho, Kel.data());
      else element(pr, e, U, nullptr, st_old, nullptr, nullptr, Kel.data(), nullptr);
      for (int d = 0; d < nf; ++d)
        for (int a = 0; a < pr.nnpe; ++a) {
          const int r = a * nf + d;
          atomic_add(&Dg[nf * (conn[e * pr.nnpe + a] - 1) + d], Kel[r * NDOF + r]);
        }
    }
  }
}

// Mass action M v into a full NF x NN field (assemble_matrix_action! with the mass callback,
// src/assemblers/MatrixAction.jl:154-238).
void orc_assemble_mass_action(int elem_kind, int nd, i64 nn, const double* X, i64 ne, const i64* conn, int nf,
                              double rho, const double* V, double* MV) {
  Problem pr = make_problem(elem_kind, nd, nn, X, ne, conn, nf, POISSON, nullptr, 0);
  const int NDOF = pr.nnpe * nf;
  for (i64 i = 0; i < nn * nf; ++i) MV[i] = 0;
#pragma omp parallel
  {
    std::vector<double> Mel(NDOF * NDOF), vel(NDOF);
#pragma omp for schedule(static)
    for (i64 e = 0; e < ne; ++e) {
      mass_element(pr, e, rho, Mel.data());
      for (int a = 0; a < pr.nnpe; ++a)
        for (int i = 0; i < nf; ++i) vel[a * nf + i] = V[nf * (conn[e * pr.nnpe + a] - 1) + i];
      for (int r = 0; r < NDOF; ++r) {
        double s = 0;
        for (int c = 0; c < NDOF; ++c) s += Mel[r * NDOF + c] * vel[c];
        const int a = r / nf, d = r % nf;
        atomic_add(&MV[nf * (conn[e * pr.nnpe + a] - 1) + d], s);
      }
    }
  }
}

// gather storage[slots] (the `coo_storage[pattern.unknown_dofs]` view, SparsityPatterns.jl:303)
void orc_gather_slots(i64 n, const i64* slots, const double* storage, double* out) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) out[i] = storage[slots[i] - 1];
}

// one J2 material-point update (for surrogate data and unit pins); grad 3x3 row-major
void orc_j2_point(const double* props, const double* grad, const double* st_old, double* st_new, double* sig, double* C) {
  j2_update(props, grad, st_old, st_new, sig, C);
}

// generic material point: P (nf*nd), A ((nf*nd)^2)
void orc_material_point(int phys, const double* props, int nprops, int nd, int nf, const double* grad, const double* Xq,
                        const double* st_old, double* st_new, double* P, double* src, double* A) {
  Material m{};
  m.kind = phys;
  for (int i = 0; i < nprops && i < 8; ++i) m.p[i] = props[i];
  material_eval(m, nd, nf, grad, Xq, st_old, st_new, P, src, A);
}

// shape functions + quadrature (for invariants tests)
int orc_ref_element(int kind, double* xi /*nq*3*/, double* w, double* N /*nq*nnpe*/, double* dN /*nq*nnpe*nd*/) {
  RefElem r = make_ref(kind);
  for (int q = 0; q < r.nq; ++q) {
    for (int d = 0; d < 3; ++d) xi[q * 3 + d] = r.xi[q][d];
    w[q] = r.w[q];
    shape(kind, r.xi[q], N + q * r.nnpe, dN + q * r.nnpe * r.nd);
  }
  return r.nq;
}

// CSR SpMV and conjugate gradients (Krylov.cg as called at src/Solvers.jl:144), FP64.
static void spmv(i64 n, const i64* rp, const i64* cv, const double* v, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) {
    double s = 0;
    for (i64 p = rp[i] - 1; p < rp[i + 1] - 1; ++p) s += v[p] * x[cv[p] - 1];
    y[i] = s;
  }
}

int orc_cg(i64 n, const i64* rowptr, const i64* colval, const double* vals, const double* b, double* x, double rtol,
           int maxit, double* final_rel) {
  std::vector<double> r(n), p(n), Ap(n);
  spmv(n, rowptr, colval, vals, x, Ap.data());
  double rr = 0, bb = 0;
  for (i64 i = 0; i < n; ++i) { r[i] = b[i] - Ap[i]; p[i] = r[i]; rr += r[i] * r[i]; bb += b[i] * b[i]; }
  double nb = std::sqrt(bb);
  if (nb == 0) nb = 1;
  int it = 0;
  while (std::sqrt(rr) > rtol * nb && it < maxit) {
    spmv(n, rowptr, colval, vals, p.data(), Ap.data());
    double pAp = 0;
    for (i64 i = 0; i < n; ++i) pAp += p[i] * Ap[i];
    double alpha = rr / pAp;
    double rr2 = 0;
    for (i64 i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; rr2 += r[i] * r[i]; }
    double beta = rr2 / rr;
    rr = rr2;
    for (i64 i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    ++it;
  }
  if (final_rel) *final_rel = std::sqrt(rr) / nb;
  return it;
}

}  // extern "C"
