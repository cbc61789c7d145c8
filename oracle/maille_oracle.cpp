// maille oracle — TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// A single-threaded FP64 CPU restatement of maille-core's 3D linear-elastic solve path
// (timothee-faget/maille v0.14.0). It is the checker the HIP library is compared against and
// the CPU baseline timed beside it. Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load it; the product library never links or calls it.
//
// Every function follows one reference function literally, including loop and summation order,
// so that it reproduces the reference's floating-point results operation for operation (the
// reference is Rust, which never contracts a*b+c into an FMA: build with -ffp-contract=off).
// Citations are file:line into crates/maille-core of the reference.
//
// Pinning (what checks each part, since the reference cannot run here):
//   shape/quadrature/Mapping3d ... closed forms: partition of unity, Kronecker, det = 6V,
//                                  grad x = I, quadrature exactness (tests/test_oracle_*.py)
//   solid_d ....................... Lame constants closed form
//   solid_stiffness ............... symmetry, 6 rigid modes in the null space, u'Ku = V e'De
//   face_pressure ................. A/3 per Tet4 vertex, A/3 per Tet10 mid-node, resultant
//   Graph / RCM / Csr pattern ..... hand-derived known-answer orders, one graph per tie-break
//                                   rule (first minimum seed, George-Liu restart, stable degree
//                                   sort, one global reverse), each failing under its planted
//                                   mistake (tests/test_rcm_known_answers.py); brute-force
//                                   restatement on random graphs, permutation + profile shrink
//   dot (8 lanes) ................. hand-derived known answers fixing the lane split, the
//                                   combination tree and the tail order (tests/test_oracle.py)
//   IC(0) ......................... tridiagonal => exact Cholesky; dense-pattern => Cholesky
//   PCG ........................... grid Laplacian recovers the chosen x; skyline agreement
//   skyline Cholesky .............. numpy dense solve; reference 4x4 known answer
//   Model::solve .................. patch test, uniaxial traction, Timoshenko cantilever,
//                                   hydrostatic cube (reference integration tests)
//   nodal stresses, edge loads .... the reference's results.rs and edge-load tests restated
//   von Mises, solid_response ..... plasticity.rs unit tests restated, tangents against finite
//                                   differences (tests/test_plasticity_oracle.py)
//   quasi-static Newton ........... tests/plasticity.rs restated: uniaxial load/unload closed
//                                   form, strain history vs the material point, elastic in one
//                                   iteration, quadratic convergence, limit load, PCG = skyline
//   large strains, follower ....... tests/finite_strain.rs restated (rigid rotations, closed
//                                   forms of both laws, follower compression, cuts), element
//                                   tangent and follower stiffness against finite differences
//   thermal, theta method ......... element closed forms; tests/transient.rs restated (Fourier
//                                   series, amplification factor, steady state, long time)
// Parity unpinned by any reference fixture (the reference ships none and cannot run here):
// exact PCG and Newton iteration counts and bitwise CSR values (element-order sums).
// The multi-core PCG variant (or_system_pcg_parallel) is a timing baseline only.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

namespace {

// ---------------------------------------------------------------- quadrature.rs
struct TetGP { double xi, eta, zeta, w; };
struct TriGP { double xi, eta, w; };

// quadrature.rs:100-105 (TETRA_1)
const TetGP TETRA_1[1] = {{0.25, 0.25, 0.25, 1.0 / 6.0}};
// quadrature.rs:108-138 (TET_A, TET_B, TETRA_4)
const double TET_A = 0.585410196624968500;
const double TET_B = 0.138196601125010500;
const TetGP TETRA_4[4] = {{TET_B, TET_B, TET_B, 1.0 / 24.0},
                          {TET_A, TET_B, TET_B, 1.0 / 24.0},
                          {TET_B, TET_A, TET_B, 1.0 / 24.0},
                          {TET_B, TET_B, TET_A, 1.0 / 24.0}};
// quadrature.rs:139-177 (TETRA_14): degree-5 rule for the Tet10 capacity
const double T14_A1 = 0.09273525031089142, T14_W1 = 0.01224884051939372;
const double T14_A2 = 0.3108859192633008, T14_W2 = 0.01878132095300283;
const double T14_B = 0.04550370412564851, T14_W3 = 0.00709100346284674;
const double T14_C1 = 1.0 - 3.0 * T14_A1, T14_C2 = 1.0 - 3.0 * T14_A2, T14_D = 0.5 - T14_B;
const TetGP TETRA_14[14] = {
    {T14_A1, T14_A1, T14_A1, T14_W1}, {T14_C1, T14_A1, T14_A1, T14_W1},
    {T14_A1, T14_C1, T14_A1, T14_W1}, {T14_A1, T14_A1, T14_C1, T14_W1},
    {T14_A2, T14_A2, T14_A2, T14_W2}, {T14_C2, T14_A2, T14_A2, T14_W2},
    {T14_A2, T14_C2, T14_A2, T14_W2}, {T14_A2, T14_A2, T14_C2, T14_W2},
    {T14_B, T14_B, T14_D, T14_W3},    {T14_B, T14_D, T14_B, T14_W3},
    {T14_D, T14_B, T14_B, T14_W3},    {T14_D, T14_D, T14_B, T14_W3},
    {T14_D, T14_B, T14_D, T14_W3},    {T14_B, T14_D, T14_D, T14_W3}};

// quadrature.rs:19-35 (TRIANGLE_3)
const TriGP TRIANGLE_3[3] = {{1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                             {2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                             {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}};
// quadrature.rs:38-54 (TRIANGLE_6, Dunavant)
const double T6A = 0.445948490915965;
const double T6WA = 0.223381589678011 / 2.0;
const double T6B = 0.091576213509771;
const double T6WB = 0.109951743655322 / 2.0;
const TriGP TRIANGLE_6[6] = {{T6A, T6A, T6WA},
                             {1.0 - 2.0 * T6A, T6A, T6WA},
                             {T6A, 1.0 - 2.0 * T6A, T6WA},
                             {T6B, T6B, T6WB},
                             {1.0 - 2.0 * T6B, T6B, T6WB},
                             {T6B, 1.0 - 2.0 * T6B, T6WB}};

// ---------------------------------------------------------------- shape.rs
// shape.rs:129 and :135
const int TET_FACES[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {2, 3, 0}};
const int TET10_EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

// shape.rs:95-111 (tet_barycentric, tet4)
void tet4_shape(double xi, double eta, double zeta, double* n, double (*dn)[3]) {
  n[0] = 1.0 - xi - eta - zeta; n[1] = xi; n[2] = eta; n[3] = zeta;
  const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 3; ++c) dn[i][c] = dl[i][c];
}

// shape.rs:138-153 (tet10)
void tet10_shape(double xi, double eta, double zeta, double* n, double (*dn)[3]) {
  const double l[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
  const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 4; ++i) {
    n[i] = l[i] * (2.0 * l[i] - 1.0);
    for (int c = 0; c < 3; ++c) dn[i][c] = (4.0 * l[i] - 1.0) * dl[i][c];
  }
  for (int k = 0; k < 6; ++k) {
    int a = TET10_EDGES[k][0], b = TET10_EDGES[k][1];
    n[4 + k] = 4.0 * l[a] * l[b];
    for (int c = 0; c < 3; ++c) dn[4 + k][c] = 4.0 * (l[a] * dl[b][c] + l[b] * dl[a][c]);
  }
}

void tet_shape(int nen, double xi, double eta, double zeta, double* n, double (*dn)[3]) {
  if (nen == 4) tet4_shape(xi, eta, zeta, n, dn);
  else tet10_shape(xi, eta, zeta, n, dn);
}

// shape.rs:10-15 (tri3) and :27-47 (tri6), values only
void tri_values(int nn, double xi, double eta, double* n) {
  if (nn == 3) {
    n[0] = 1.0 - xi - eta; n[1] = xi; n[2] = eta;
    return;
  }
  double l1 = 1.0 - xi - eta, l2 = xi, l3 = eta;
  n[0] = l1 * (2.0 * l1 - 1.0);
  n[1] = l2 * (2.0 * l2 - 1.0);
  n[2] = l3 * (2.0 * l3 - 1.0);
  n[3] = 4.0 * l1 * l2;
  n[4] = 4.0 * l2 * l3;
  n[5] = 4.0 * l3 * l1;
}

// shape.rs:157-190 (Mapping3d::new)
double mapping3d(int nen, const double (*x)[3], const double (*dn)[3], double (*dn_dx)[3]) {
  double j[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int k = 0; k < nen; ++k)
    for (int i = 0; i < 3; ++i)
      for (int c = 0; c < 3; ++c) j[i][c] += dn[k][i] * x[k][c];
  auto cof = [&](int r, int c) {
    int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    return j[r1][c1] * j[r2][c2] - j[r1][c2] * j[r2][c1];
  };
  double det = 0.0;
  for (int c = 0; c < 3; ++c) det += j[0][c] * cof(0, c);
  double inv[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) inv[r][c] = cof(c, r) / det;
  for (int k = 0; k < nen; ++k)
    for (int r = 0; r < 3; ++r) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += inv[r][c] * dn[k][c];
      dn_dx[k][r] = s;
    }
  return det;
}

// ---------------------------------------------------------------- element.rs
// element.rs:193-206 (solid_d)
void solid_d(double young, double nu, double d[6][6]) {
  double c = young / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double g = young / (2.0 * (1.0 + nu));
  for (int i = 0; i < 6; ++i)
    for (int k = 0; k < 6; ++k) d[i][k] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) d[i][k] = (i == k) ? c * (1.0 - nu) : c * nu;
  for (int i = 3; i < 6; ++i) d[i][i] = g;
}

// element.rs:209-220 (solid_strain_displacement): column i of B, Voigt (xx,yy,zz,yz,xz,xy)
void b_column(const double (*g)[3], int i, double col[6]) {
  const double gx = g[i / 3][0], gy = g[i / 3][1], gz = g[i / 3][2];
  switch (i % 3) {
    case 0: col[0] = gx; col[1] = 0; col[2] = 0; col[3] = 0; col[4] = gz; col[5] = gy; break;
    case 1: col[0] = 0; col[1] = gy; col[2] = 0; col[3] = gz; col[4] = 0; col[5] = gx; break;
    default: col[0] = 0; col[1] = 0; col[2] = gz; col[3] = gy; col[4] = gx; col[5] = 0; break;
  }
}

const TetGP* rule_of(int nen, int* ngp) {
  // assembly.rs:204-230 (SolidKernel::tet4 -> TETRA_1, ::tet10 -> TETRA_4)
  if (nen == 4) { *ngp = 1; return TETRA_1; }
  *ngp = 4;
  return TETRA_4;
}

// element.rs:223-243 (solid_stiffness); ke is m x m row-major, ke[j*m+i] = ke[j][i]
void solid_stiffness(int nen, const double (*x)[3], const double d[6][6], double* ke) {
  const int m = 3 * nen;
  int ngp;
  const TetGP* rule = rule_of(nen, &ngp);
  std::fill(ke, ke + m * m, 0.0);
  double n[10], dn[10][3], g[10][3];
  for (int q = 0; q < ngp; ++q) {
    tet_shape(nen, rule[q].xi, rule[q].eta, rule[q].zeta, n, dn);
    double det = mapping3d(nen, x, dn, g);
    double factor = rule[q].w * std::fabs(det);
    for (int i = 0; i < m; ++i) {
      double bi[6], db[6];
      b_column(g, i, bi);
      for (int r = 0; r < 6; ++r) {
        double s = 0.0;
        for (int c = 0; c < 6; ++c) s += d[r][c] * bi[c];
        db[r] = s;
      }
      for (int jj = 0; jj < m; ++jj) {
        double bj[6];
        b_column(g, jj, bj);
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += bj[r] * db[r];
        ke[jj * m + i] += factor * s;
      }
    }
  }
}

// element.rs:270-302 (solid_stresses) without thermal strain; out per GP:
// x, y, z, sxx, syy, szz, tyz, txz, txy (Voigt order after the position)
void solid_stresses(int nen, const double (*x)[3], const double d[6][6], const double* ue,
                    double* out) {
  int ngp;
  const TetGP* rule = rule_of(nen, &ngp);
  const int m = 3 * nen;
  double n[10], dn[10][3], g[10][3];
  for (int q = 0; q < ngp; ++q) {
    tet_shape(nen, rule[q].xi, rule[q].eta, rule[q].zeta, n, dn);
    mapping3d(nen, x, dn, g);
    double eps[6];
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int k = 0; k < m; ++k) {
        double col[6];
        b_column(g, k, col);
        s += col[r] * ue[k];
      }
      eps[r] = s;  // minus 0.0 (no thermal strain): identity
    }
    double* o = out + 9 * q;
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int a = 0; a < nen; ++a) s += n[a] * x[a][k];
      o[k] = s;
    }
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int c = 0; c < 6; ++c) s += d[r][c] * eps[c];
      o[3 + r] = s;
    }
  }
}

// ---------------------------------------------------------------- plasticity.rs
// plasticity.rs:11 (YIELD_TOLERANCE)
const double YIELD_TOLERANCE = 1e-10;

// plasticity.rs:14-21 (PlasticState): plastic strain (Voigt, engineering shears) and p
struct PlasticState { double ep[6]; double p; };

// plasticity.rs:115-126 (contract): shears count twice
double contract6(const double* a, const double* b) {
  double s = 0.0;
  for (int k = 0; k < 6; ++k) s += k < 3 ? a[k] * b[k] : 2.0 * a[k] * b[k];
  return s;
}

// plasticity.rs:45-112 (VonMises::update): closed-form radial return from `prev`, stress, end
// state, consistent tangent (6x6) and whether the point yielded.
bool von_mises_update(double young, double poisson, double yield0, double hardening,
                      const double strain[6], const PlasticState& prev, double stress[6],
                      PlasticState& state, double tangent[6][6]) {
  const double g = young / (2.0 * (1.0 + poisson));
  const double bulk = young / (3.0 * (1.0 - 2.0 * poisson));
  double elastic[6];
  for (int k = 0; k < 6; ++k) elastic[k] = strain[k] - prev.ep[k];
  const double volume = elastic[0] + elastic[1] + elastic[2];
  double trial[6];
  for (int k = 0; k < 6; ++k)
    trial[k] = k < 3 ? 2.0 * g * (elastic[k] - volume / 3.0) : g * elastic[k];
  const double root_3_2 = std::sqrt(1.5);
  const double norm = std::sqrt(contract6(trial, trial));
  const double q = root_3_2 * norm;
  const double yield_stress = yield0 + hardening * prev.p;
  const double excess = q - yield_stress;
  double theta = 1.0, theta_bar = 0.0, dp = 0.0;
  if (excess > YIELD_TOLERANCE * yield_stress) {
    dp = excess / (3.0 * g + hardening);
    theta = 1.0 - 3.0 * g * dp / q;
    theta_bar = 3.0 * g / (3.0 * g + hardening) - (1.0 - theta);
  }
  double n[6];
  for (int k = 0; k < 6; ++k) n[k] = dp > 0.0 ? trial[k] / norm : 0.0;
  state = prev;
  for (int k = 0; k < 6; ++k) {
    const double tensor = root_3_2 * dp * n[k];
    state.ep[k] += k < 3 ? tensor : 2.0 * tensor;
  }
  state.p += dp;
  for (int k = 0; k < 6; ++k) stress[k] = theta * trial[k] + (k < 3 ? bulk * volume : 0.0);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      const double delta = i == j ? 1.0 : 0.0;
      double deviatoric = 0.0;
      if (i < 3 && j < 3) deviatoric = delta - 1.0 / 3.0;
      else if (i >= 3 && j >= 3 && i == j) deviatoric = 0.5;
      const double volumetric = (i < 3 && j < 3) ? bulk : 0.0;
      tangent[i][j] =
          volumetric + 2.0 * g * theta * deviatoric - 2.0 * g * theta_bar * n[i] * n[j];
    }
  return dp > 0.0;
}

// element.rs:338-358 (solid_gauss) + :395-452 (solid_response, elastoplastic_response) for one
// tetrahedron: internal force fint[m], tangent tan[j*m+i] = tangent[j][i] (row-major), Gauss-point
// stresses out[q*9..] = x, y, z, s0..s5 (Voigt), end states, yielding.
bool solid_response(int nen, const double (*x)[3], double young, double poisson, double yield0,
                    double hardening, const double* ue, const PlasticState* prev, double* fint,
                    double* tan, double* out, PlasticState* states) {
  const int m = 3 * nen;
  int ngp;
  const TetGP* rule = rule_of(nen, &ngp);
  std::fill(fint, fint + m, 0.0);
  std::fill(tan, tan + m * m, 0.0);
  bool yielding = false;
  double n[10], dn[10][3], g[10][3];
  for (int q = 0; q < ngp; ++q) {
    tet_shape(nen, rule[q].xi, rule[q].eta, rule[q].zeta, n, dn);
    const double det = mapping3d(nen, x, dn, g);
    const double weight = rule[q].w * std::fabs(det);
    double pos[3];
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int a = 0; a < nen; ++a) s += n[a] * x[a][k];
      pos[k] = s;
    }
    double strain[6];
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int k = 0; k < m; ++k) {
        double col[6];
        b_column(g, k, col);
        s += col[r] * ue[k];
      }
      strain[r] = s;
    }
    double stress[6], c[6][6];
    yielding |= von_mises_update(young, poisson, yield0, hardening, strain, prev[q], stress,
                                 states[q], c);
    for (int i = 0; i < m; ++i) {
      double bi[6], cb[6];
      b_column(g, i, bi);
      double f = 0.0;
      for (int r = 0; r < 6; ++r) f += bi[r] * stress[r];
      fint[i] += weight * f;
      for (int r = 0; r < 6; ++r) {
        double s = 0.0;
        for (int cc = 0; cc < 6; ++cc) s += c[r][cc] * bi[cc];
        cb[r] = s;
      }
      for (int j = 0; j < m; ++j) {
        double bj[6];
        b_column(g, j, bj);
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += bj[r] * cb[r];
        tan[j * m + i] += weight * s;
      }
    }
    double* o = out + 9 * q;
    for (int k = 0; k < 3; ++k) o[k] = pos[k];
    for (int k = 0; k < 6; ++k) o[3 + k] = stress[k];
  }
  return yielding;
}

// mesh.rs:184-198 (face_positions, tet part)
int face_positions(int nen, int face, int* pos) {
  const int* f = TET_FACES[face];
  pos[0] = f[0]; pos[1] = f[1]; pos[2] = f[2];
  if (nen == 4) return 3;
  auto mid = [](int a, int b) {
    for (int k = 0; k < 6; ++k)
      if ((TET10_EDGES[k][0] == a && TET10_EDGES[k][1] == b) ||
          (TET10_EDGES[k][0] == b && TET10_EDGES[k][1] == a))
        return 4 + k;
    return 4;
  };
  pos[3] = mid(f[0], f[1]); pos[4] = mid(f[1], f[2]); pos[5] = mid(f[2], f[0]);
  return 6;
}

// assembly.rs:527-593 (face_rule, tet branch) + :650-674 (face_pressure)
// forces: 3 per face node, in face_positions order. Returns the number of face nodes.
int face_pressure(int nen, const double (*xe)[3], int face, double value, int* pos,
                  double* forces) {
  int nn = face_positions(nen, face, pos);
  double c[6][3];
  for (int a = 0; a < nn; ++a)
    for (int k = 0; k < 3; ++k) c[a][k] = xe[pos[a]][k];
  double u[3], v[3];
  for (int k = 0; k < 3; ++k) { u[k] = c[1][k] - c[0][k]; v[k] = c[2][k] - c[0][k]; }
  double cross[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                     u[0] * v[1] - u[1] * v[0]};
  double sq = 0.0;
  for (int k = 0; k < 3; ++k) sq += cross[k] * cross[k];
  double twice_area = std::sqrt(sq);
  int ngp = (nn == 3) ? 3 : 6;
  const TriGP* rule = (nn == 3) ? TRIANGLE_3 : TRIANGLE_6;
  double outward[3];
  for (int k = 0; k < 3; ++k)
    outward[k] = (c[0][k] + c[1][k] + c[2][k]) / 3.0 -
                 (xe[0][k] + xe[1][k] + xe[2][k] + xe[3][k]) / 4.0;
  double dot = 0.0;
  for (int k = 0; k < 3; ++k) dot += cross[k] * outward[k];
  double sign = dot < 0.0 ? -1.0 : 1.0;
  double normal[3];
  for (int k = 0; k < 3; ++k) normal[k] = sign * cross[k] / twice_area;
  for (int a = 0; a < 3 * nn; ++a) forces[a] = 0.0;
  for (int q = 0; q < ngp; ++q) {
    double w = rule[q].w * twice_area;
    double nv[6];
    tri_values(nn, rule[q].xi, rule[q].eta, nv);
    for (int a = 0; a < nn; ++a)
      for (int k = 0; k < 3; ++k) forces[3 * a + k] -= value * w * nv[a] * normal[k];
  }
  return nn;
}

// ---------------------------------------------------------------- finite_strain.rs
typedef double Mat3[3][3];

// finite_strain.rs:89-117 (mul, transpose, det, inverse)
void mat_mul(const Mat3 a, const Mat3 b, Mat3 out) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += a[i][k] * b[k][j];
      out[i][j] = s;
    }
}
double mat_det(const Mat3 a) {
  return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
         a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}
void mat_inverse(const Mat3 a, Mat3 out) {
  const double d = mat_det(a);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      // cofactor(j, i)
      const int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
      out[i][j] = (a[r0][c0] * a[r1][c1] - a[r0][c1] * a[r1][c0]) / d;
    }
}
inline double kd(int a, int b) { return a == b ? 1.0 : 0.0; }

// finite_strain.rs:39-86 (LargeStrainLaw::evaluate): second Piola-Kirchhoff stress and the
// material modulus; neo = Neo-Hookean, else Saint-Venant-Kirchhoff
void law_evaluate(bool neo, double lambda, double mu, const Mat3 f, Mat3 s, double* modulus) {
  Mat3 ft, c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ft[i][j] = f[j][i];
  mat_mul(ft, f, c);
  if (!neo) {
    Mat3 e;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) e[i][j] = 0.5 * (c[i][j] - kd(i, j));
    const double trace = e[0][0] + e[1][1] + e[2][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) s[i][j] = 2.0 * mu * e[i][j] + lambda * trace * kd(i, j);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l)
            modulus[((i * 3 + j) * 3 + k) * 3 + l] =
                lambda * kd(i, j) * kd(k, l) + mu * (kd(i, k) * kd(j, l) + kd(i, l) * kd(j, k));
    return;
  }
  const double log_j = std::log(mat_det(f));
  Mat3 inv;
  mat_inverse(c, inv);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s[i][j] = mu * (kd(i, j) - inv[i][j]) + lambda * log_j * inv[i][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          modulus[((i * 3 + j) * 3 + k) * 3 + l] =
              lambda * inv[i][j] * inv[k][l] +
              (mu - lambda * log_j) * (inv[i][k] * inv[j][l] + inv[i][l] * inv[j][k]);
}

// finite_strain.rs:119-214 (large_strain_response) for a tetrahedron on its element rule
// (element.rs:338-358): internal force, tangent tan[row*m + col], Cauchy stresses at the current
// Gauss-point positions out[q*9..] = x, y, z, sxx, syy, szz, tyz, txz, txy.
void large_strain_response(int nen, const double (*x)[3], bool neo, double lambda, double mu,
                           const double* ue, double* fint, double* tan, double* out) {
  const int m = 3 * nen;
  int ngp;
  const TetGP* rule = rule_of(nen, &ngp);
  std::fill(fint, fint + m, 0.0);
  std::fill(tan, tan + m * m, 0.0);
  double n[10], dn[10][3], g[10][3], modulus[81];
  for (int q = 0; q < ngp; ++q) {
    tet_shape(nen, rule[q].xi, rule[q].eta, rule[q].zeta, n, dn);
    const double weight = rule[q].w * std::fabs(mapping3d(nen, x, dn, g));
    double pos[3];
    for (int c = 0; c < 3; ++c) {
      double sm = 0.0;
      for (int a = 0; a < nen; ++a) sm += n[a] * x[a][c];
      pos[c] = sm;
    }
    Mat3 f;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) f[i][j] = i == j ? 1.0 : 0.0;
    for (int a = 0; a < nen; ++a)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) f[i][j] += ue[3 * a + i] * g[a][j];
    Mat3 ft, s, pk1;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) ft[i][j] = f[j][i];
    law_evaluate(neo, lambda, mu, f, s, modulus);
    mat_mul(f, s, pk1);
    double a4[81];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            double material = 0.0;
            for (int ii = 0; ii < 3; ++ii)
              for (int kk = 0; kk < 3; ++kk)
                material += f[i][ii] * modulus[((ii * 3 + j) * 3 + kk) * 3 + l] * f[k][kk];
            const double geometric = i == k ? s[j][l] : 0.0;
            a4[((i * 3 + j) * 3 + k) * 3 + l] = geometric + material;
          }
    for (int na = 0; na < nen; ++na)
      for (int i = 0; i < 3; ++i) {
        const int row = 3 * na + i;
        double fsum = 0.0;
        for (int j = 0; j < 3; ++j) fsum += pk1[i][j] * g[na][j];
        fint[row] += weight * fsum;
        for (int nb = 0; nb < nen; ++nb)
          for (int k = 0; k < 3; ++k) {
            double v = 0.0;
            for (int j = 0; j < 3; ++j)
              for (int l = 0; l < 3; ++l) v += g[na][j] * a4[((i * 3 + j) * 3 + k) * 3 + l] * g[nb][l];
            tan[row * m + 3 * nb + k] += weight * v;
          }
      }
    const double jdet = mat_det(f);
    Mat3 sigma;
    mat_mul(pk1, ft, sigma);
    double* o = out + 9 * q;
    for (int c = 0; c < 3; ++c) {
      double sm = 0.0;
      for (int a = 0; a < nen; ++a) sm += n[a] * ue[3 * a + c];
      o[c] = pos[c] + sm;
    }
    o[3] = sigma[0][0] / jdet;
    o[4] = sigma[1][1] / jdet;
    o[5] = sigma[2][2] / jdet;
    o[6] = sigma[1][2] / jdet;
    o[7] = sigma[0][2] / jdet;
    o[8] = sigma[0][1] / jdet;
  }
}

// assembly.rs:727-730 (levi_civita)
double levi_civita(int i, int j, int k) {
  const double a = i, b = j, c = k;
  return (a - b) * (b - c) * (c - a) / 2.0;
}

// assembly.rs:732-862 (follower_pressure), tetrahedron face: forces[3 m] per unit lambda and
// stiffness[3m x 3m] of the pressure `value` on `face` at the displacements u (global).
// Returns the face node count m; pos[] their positions in the element.
int follower_pressure(int nen, const i64* enodes, const double* xyz, int face, double value,
                      const double* u, int* pos, double* forces, double* stiffness) {
  const int m = face_positions(nen, face, pos);
  double ref[6][3], cur[6][3], xe[10][3];
  for (int a = 0; a < nen; ++a)
    for (int c = 0; c < 3; ++c) xe[a][c] = xyz[3 * enodes[a] + c];
  for (int a = 0; a < m; ++a)
    for (int c = 0; c < 3; ++c) {
      ref[a][c] = xe[pos[a]][c];
      cur[a][c] = ref[a][c] + u[3 * enodes[pos[a]] + c];
    }
  // face_rule's outward normal (assembly.rs:555-584) on the reference coordinates
  double uu[3], vv[3];
  for (int k = 0; k < 3; ++k) { uu[k] = ref[1][k] - ref[0][k]; vv[k] = ref[2][k] - ref[0][k]; }
  double cross[3] = {uu[1] * vv[2] - uu[2] * vv[1], uu[2] * vv[0] - uu[0] * vv[2],
                     uu[0] * vv[1] - uu[1] * vv[0]};
  double sq = 0.0;
  for (int k = 0; k < 3; ++k) sq += cross[k] * cross[k];
  const double twice_area = std::sqrt(sq);
  double outward[3];
  for (int k = 0; k < 3; ++k)
    outward[k] = (ref[0][k] + ref[1][k] + ref[2][k]) / 3.0 -
                 (xe[0][k] + xe[1][k] + xe[2][k] + xe[3][k]) / 4.0;
  double dotn = 0.0;
  for (int k = 0; k < 3; ++k) dotn += cross[k] * outward[k];
  const double nsign = dotn < 0.0 ? -1.0 : 1.0;
  double normal[3];
  for (int k = 0; k < 3; ++k) normal[k] = nsign * cross[k] / twice_area;
  const int ngp = m == 3 ? 3 : 6;
  const TriGP* rule = m == 3 ? TRIANGLE_3 : TRIANGLE_6;
  double nv[6][6], dn[6][6][2];
  for (int q = 0; q < ngp; ++q) {
    const double xi = rule[q].xi, eta = rule[q].eta;
    if (m == 3) {
      nv[q][0] = 1.0 - xi - eta; nv[q][1] = xi; nv[q][2] = eta;
      const double d[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
      for (int a = 0; a < 3; ++a) { dn[q][a][0] = d[a][0]; dn[q][a][1] = d[a][1]; }
    } else {
      const double l1 = 1.0 - xi - eta, l2 = xi, l3 = eta;
      nv[q][0] = l1 * (2.0 * l1 - 1.0);
      nv[q][1] = l2 * (2.0 * l2 - 1.0);
      nv[q][2] = l3 * (2.0 * l3 - 1.0);
      nv[q][3] = 4.0 * l1 * l2;
      nv[q][4] = 4.0 * l2 * l3;
      nv[q][5] = 4.0 * l3 * l1;
      const double d[6][2] = {{1.0 - 4.0 * l1, 1.0 - 4.0 * l1}, {4.0 * l2 - 1.0, 0.0},
                              {0.0, 4.0 * l3 - 1.0},           {4.0 * (l1 - l2), -4.0 * l2},
                              {4.0 * l3, 4.0 * l2},            {-4.0 * l3, 4.0 * (l1 - l3)}};
      for (int a = 0; a < 6; ++a) { dn[q][a][0] = d[a][0]; dn[q][a][1] = d[a][1]; }
    }
  }
  auto tangents = [&](const double (*xx)[3], int q, double* a, double* b) {
    for (int c = 0; c < 3; ++c) {
      double s0 = 0.0, s1 = 0.0;
      for (int k = 0; k < m; ++k) s0 += dn[q][k][0] * xx[k][c];
      for (int k = 0; k < m; ++k) s1 += dn[q][k][1] * xx[k][c];
      a[c] = s0;
      b[c] = s1;
    }
  };
  auto area_vector = [](const double* a, const double* b, double* out) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
  };
  double a0[3], b0[3], av[3];
  tangents(ref, 0, a0, b0);
  area_vector(a0, b0, av);
  double along = 0.0;
  for (int c = 0; c < 3; ++c) along += av[c] * normal[c];
  const double sign = along < 0.0 ? -1.0 : 1.0;
  std::fill(forces, forces + 3 * m, 0.0);
  std::fill(stiffness, stiffness + 9 * m * m, 0.0);
  for (int q = 0; q < ngp; ++q) {
    double xi[3], eta[3], area[3];
    tangents(cur, q, xi, eta);
    area_vector(xi, eta, area);
    const double w = -value * sign * rule[q].w;
    for (int a = 0; a < m; ++a)
      for (int c = 0; c < 3; ++c) forces[3 * a + c] += w * nv[q][a] * area[c];
    for (int b = 0; b < m; ++b)
      for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) {
          double d = 0.0;
          for (int l = 0; l < 3; ++l)
            d += levi_civita(c, k, l) * dn[q][b][0] * eta[l] + levi_civita(c, l, k) * xi[l] * dn[q][b][1];
          for (int a = 0; a < m; ++a) stiffness[(3 * a + c) * 3 * m + 3 * b + k] += w * nv[q][a] * d;
        }
  }
  return m;
}

// ---------------------------------------------------------------- thermal elements
// element.rs:450-462 (conduction) on the element's rule (assembly.rs:392-403):
// ke[i*nen + j] += w|det| k (g_i . g_j)
void conduction(int nen, const double (*x)[3], double k, double* ke) {
  int ngp;
  const TetGP* rule = rule_of(nen, &ngp);
  std::fill(ke, ke + nen * nen, 0.0);
  double n[10], dn[10][3], g[10][3];
  for (int q = 0; q < ngp; ++q) {
    tet_shape(nen, rule[q].xi, rule[q].eta, rule[q].zeta, n, dn);
    const double weight = rule[q].w * std::fabs(mapping3d(nen, x, dn, g));
    for (int i = 0; i < nen; ++i)
      for (int j = 0; j < nen; ++j) {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) s += g[i][c] * g[j][c];
        ke[i * nen + j] += weight * k * s;
      }
  }
}

// element.rs:465-477 (capacity) with TETRA_4 (Tet4) or TETRA_14 (Tet10), assembly.rs:413-431
void capacity(int nen, const double (*x)[3], double rho_c, double* ce) {
  const TetGP* rule = nen == 4 ? TETRA_4 : TETRA_14;
  const int ngp = nen == 4 ? 4 : 14;
  std::fill(ce, ce + nen * nen, 0.0);
  double n[10], dn[10][3], g[10][3];
  for (int q = 0; q < ngp; ++q) {
    tet_shape(nen, rule[q].xi, rule[q].eta, rule[q].zeta, n, dn);
    const double weight = rule[q].w * std::fabs(mapping3d(nen, x, dn, g));
    for (int i = 0; i < nen; ++i)
      for (int j = 0; j < nen; ++j) ce[i * nen + j] += weight * rho_c * n[i] * n[j];
  }
}

// assembly.rs:465-484 (element_heat_source, nodal_source): f_i += w|det| N_i q
void nodal_source(int nen, const double (*x)[3], double q, double* f) {
  int ngp;
  const TetGP* rule = rule_of(nen, &ngp);
  std::fill(f, f + nen, 0.0);
  double n[10], dn[10][3], g[10][3];
  for (int p = 0; p < ngp; ++p) {
    tet_shape(nen, rule[p].xi, rule[p].eta, rule[p].zeta, n, dn);
    const double weight = rule[p].w * std::fabs(mapping3d(nen, x, dn, g));
    for (int i = 0; i < nen; ++i) f[i] += weight * n[i] * q;
  }
}

// assembly.rs:598-624 (face_convection) on the face rule of assembly.rs:527-593: matrix
// (face nodes x face nodes, row-major) += film w N_a N_b, load += film sink w N_a
int face_convection(int nen, const double (*xe)[3], int face, double film, double sink, int* pos,
                    double* matrix, double* load) {
  int nn = face_positions(nen, face, pos);
  double c[6][3];
  for (int a = 0; a < nn; ++a)
    for (int k = 0; k < 3; ++k) c[a][k] = xe[pos[a]][k];
  double u[3], v[3];
  for (int k = 0; k < 3; ++k) { u[k] = c[1][k] - c[0][k]; v[k] = c[2][k] - c[0][k]; }
  double cross[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                     u[0] * v[1] - u[1] * v[0]};
  double sq = 0.0;
  for (int k = 0; k < 3; ++k) sq += cross[k] * cross[k];
  const double twice_area = std::sqrt(sq);
  const int ngp = (nn == 3) ? 3 : 6;
  const TriGP* rule = (nn == 3) ? TRIANGLE_3 : TRIANGLE_6;
  std::fill(matrix, matrix + nn * nn, 0.0);
  std::fill(load, load + nn, 0.0);
  for (int q = 0; q < ngp; ++q) {
    const double w = rule[q].w * twice_area;
    double nv[6];
    tri_values(nn, rule[q].xi, rule[q].eta, nv);
    for (int a = 0; a < nn; ++a) {
      load[a] += film * sink * w * nv[a];
      for (int b = 0; b < nn; ++b) matrix[a * nn + b] += film * w * nv[a] * nv[b];
    }
  }
  return nn;
}

// element.rs:478-499 (heat_fluxes): per Gauss point x, y, z, qx, qy, qz with q = -k grad T
void heat_fluxes(int nen, const double (*x)[3], double k, const double* te, double* out) {
  int ngp;
  const TetGP* rule = rule_of(nen, &ngp);
  double n[10], dn[10][3], g[10][3];
  for (int q = 0; q < ngp; ++q) {
    tet_shape(nen, rule[q].xi, rule[q].eta, rule[q].zeta, n, dn);
    mapping3d(nen, x, dn, g);
    double* o = out + 6 * q;
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int a = 0; a < nen; ++a) s += n[a] * x[a][c];
      o[c] = s;
    }
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int a = 0; a < nen; ++a) s += g[a][c] * te[a];
      o[3 + c] = -k * s;
    }
  }
}

// ---------------------------------------------------------------- solve.rs: dot
// solve.rs:100-110: 8 lanes over chunks of 8, tail summed separately
double dot8(const double* a, const double* b, i64 n) {
  i64 chunks = n / 8;
  double tail = 0.0;
  for (i64 k = chunks * 8; k < n; ++k) tail += a[k] * b[k];
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (i64 c = 0; c < chunks; ++c)
    for (int k = 0; k < 8; ++k) s[k] += a[8 * c + k] * b[8 * c + k];
  return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])) + tail;
}

// ---------------------------------------------------------------- iterative.rs: Graph
struct Graph {
  std::vector<i64> offsets, targets;
  i64 n() const { return (i64)offsets.size() - 1; }
  i64 degree(i64 v) const { return offsets[v + 1] - offsets[v]; }
};

// iterative.rs:38-76 (Graph::from_elements); elements given as CSR (eoff, eitems)
Graph graph_from_elements(i64 n, i64 ne, const i64* eoff, const i64* eit) {
  std::vector<i64> incident(n + 1, 0);
  for (i64 e = 0; e < ne; ++e)
    for (i64 p = eoff[e]; p < eoff[e + 1]; ++p) incident[eit[p] + 1] += 1;
  for (i64 v = 0; v < n; ++v) incident[v + 1] += incident[v];
  std::vector<i64> cursor(incident);
  std::vector<i64> of_vertex(incident[n]);
  for (i64 e = 0; e < ne; ++e)
    for (i64 p = eoff[e]; p < eoff[e + 1]; ++p) of_vertex[cursor[eit[p]]++] = e;
  Graph g;
  g.offsets.reserve(n + 1);
  g.offsets.push_back(0);
  std::vector<i64> mark(n, -1);
  for (i64 v = 0; v < n; ++v) {
    mark[v] = v;
    for (i64 q = incident[v]; q < incident[v + 1]; ++q) {
      i64 e = of_vertex[q];
      for (i64 p = eoff[e]; p < eoff[e + 1]; ++p) {
        i64 w = eit[p];
        if (mark[w] != v) { mark[w] = v; g.targets.push_back(w); }
      }
    }
    g.offsets.push_back((i64)g.targets.size());
  }
  return g;
}

// solve.rs:248-270 (bfs_levels)
std::vector<std::vector<i64>> bfs_levels(const Graph& g, i64 start, const std::vector<char>& excl) {
  std::vector<char> seen(excl);
  seen[start] = 1;
  std::vector<std::vector<i64>> levels;
  levels.push_back({start});
  for (;;) {
    std::vector<i64> next;
    for (i64 v : levels.back())
      for (i64 p = g.offsets[v]; p < g.offsets[v + 1]; ++p) {
        i64 w = g.targets[p];
        if (!seen[w]) { seen[w] = 1; next.push_back(w); }
      }
    if (next.empty()) return levels;
    levels.push_back(std::move(next));
  }
}

// solve.rs:200-245 (reverse_cuthill_mckee): order[new] = old
std::vector<i64> rcm(const Graph& g) {
  const i64 n = g.n();
  std::vector<char> visited(n, 0);
  std::vector<i64> order;
  order.reserve(n);
  for (;;) {
    // first unvisited vertex of minimum degree (min_by_key keeps the first minimum)
    i64 seed = -1;
    for (i64 v = 0; v < n; ++v)
      if (!visited[v] && (seed < 0 || g.degree(v) < g.degree(seed))) seed = v;
    if (seed < 0) break;
    i64 start = seed;
    auto levels = bfs_levels(g, start, visited);
    for (;;) {
      const auto& last = levels.back();
      i64 far = -1;
      for (i64 v : last)
        if (far < 0 || g.degree(v) < g.degree(far)) far = v;
      if (far < 0) break;
      auto far_levels = bfs_levels(g, far, visited);
      if (far_levels.size() <= levels.size()) break;
      start = far;
      levels = std::move(far_levels);
    }
    i64 head = (i64)order.size();
    visited[start] = 1;
    order.push_back(start);
    std::vector<i64> next;
    while (head < (i64)order.size()) {
      i64 v = order[head++];
      next.clear();
      for (i64 p = g.offsets[v]; p < g.offsets[v + 1]; ++p)
        if (!visited[g.targets[p]]) next.push_back(g.targets[p]);
      std::stable_sort(next.begin(), next.end(),
                       [&](i64 a, i64 b) { return g.degree(a) < g.degree(b); });
      for (i64 w : next)
        if (!visited[w]) { visited[w] = 1; order.push_back(w); }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// ---------------------------------------------------------------- iterative.rs: Csr
struct Csr {
  std::vector<i64> offsets, cols;
  std::vector<double> values;
  i64 n() const { return (i64)offsets.size() - 1; }

  // iterative.rs:91-114 (Csr::zeros)
  static Csr zeros(const Graph& g, const std::vector<i64>& order, const std::vector<i64>& rank) {
    Csr a;
    a.offsets.reserve(order.size() + 1);
    a.offsets.push_back(0);
    for (i64 r = 0; r < (i64)order.size(); ++r) {
      i64 v = order[r];
      size_t start = a.cols.size();
      for (i64 p = g.offsets[v]; p < g.offsets[v + 1]; ++p) {
        i64 c = rank[g.targets[p]];
        if (c < r) a.cols.push_back(c);
      }
      std::sort(a.cols.begin() + start, a.cols.end());
      a.cols.push_back(r);
      a.offsets.push_back((i64)a.cols.size());
    }
    a.values.assign(a.cols.size(), 0.0);
    return a;
  }

  // iterative.rs:121-128 (Csr::add); returns false where the reference panics
  bool add(i64 i, i64 j, double v) {
    if (j > i) std::swap(i, j);
    auto b = cols.begin() + offsets[i], e = cols.begin() + offsets[i + 1];
    auto it = std::lower_bound(b, e, j);
    if (it == e || *it != j) return false;
    values[it - cols.begin()] += v;
    return true;
  }

  // iterative.rs:131-142 (multiply)
  void multiply(const double* x, double* y) const {
    const i64 nn = n();
    std::fill(y, y + nn, 0.0);
    for (i64 i = 0; i < nn; ++i) {
      i64 start = offsets[i], end = offsets[i + 1] - 1;
      double s = values[end] * x[i];
      for (i64 p = start; p < end; ++p) {
        s += values[p] * x[cols[p]];
        y[cols[p]] += values[p] * x[i];
      }
      y[i] += s;
    }
  }

  // iterative.rs:146-170 (incomplete_cholesky); empty on failure
  bool incomplete_cholesky(double shift, std::vector<double>& l) const {
    const i64 nn = n();
    l.assign(values.size(), 0.0);
    std::vector<double> w(nn, 0.0);
    for (i64 i = 0; i < nn; ++i) {
      i64 start = offsets[i], diag = offsets[i + 1] - 1;
      for (i64 p = start; p < diag; ++p) {
        i64 j = cols[p];
        i64 js = offsets[j], jd = offsets[j + 1] - 1;
        double s = 0.0;
        for (i64 q = js; q < jd; ++q) s += l[q] * w[cols[q]];
        l[p] = (values[p] - s) / l[jd];
        w[j] = l[p];
      }
      double squares = 0.0;
      for (i64 p = start; p < diag; ++p) squares += l[p] * l[p];
      double pivot = values[diag] * (1.0 + shift) - squares;
      if (std::isnan(pivot) || pivot <= 1e-12 * values[diag]) return false;
      l[diag] = std::sqrt(pivot);
      for (i64 p = start; p < diag; ++p) w[cols[p]] = 0.0;
    }
    return true;
  }

  // iterative.rs:173-188 (precondition)
  void precondition(const std::vector<double>& l, const double* r, double* z) const {
    const i64 nn = n();
    std::copy(r, r + nn, z);
    for (i64 i = 0; i < nn; ++i) {
      i64 start = offsets[i], diag = offsets[i + 1] - 1;
      double s = 0.0;
      for (i64 p = start; p < diag; ++p) s += l[p] * z[cols[p]];
      z[i] = (z[i] - s) / l[diag];
    }
    for (i64 i = nn - 1; i >= 0; --i) {
      i64 start = offsets[i], diag = offsets[i + 1] - 1;
      z[i] /= l[diag];
      for (i64 p = start; p < diag; ++p) z[cols[p]] -= l[p] * z[i];
    }
  }
};

// iterative.rs:9-14
const double TOLERANCE = 1e-12;
const i64 MAX_ITERATIONS = 10000;
const double SHIFTS[6] = {0.0, 1e-3, 1e-2, 1e-1, 1.0, 10.0};

enum Status { OK = 0, SINGULAR = 1, NOT_CONVERGED = 2, INVALID = 3, OUT_OF_PATTERN = 4 };

// iterative.rs:201-207 (Pcg::new): first shift of the ladder that factors
int pcg_new(const Csr& a, std::vector<double>& l, int* shift_index) {
  for (int s = 0; s < 6; ++s)
    if (a.incomplete_cholesky(SHIFTS[s], l)) {
      if (shift_index) *shift_index = s;
      return OK;
    }
  return SINGULAR;
}

// iterative.rs:210-248 (Pcg::solve), with the tolerance and iteration cap as arguments
int pcg_solve(const Csr& a, const std::vector<double>& l, const double* b, double* x,
              double tol, i64 maxit, i64* iterations, double* residual) {
  const i64 n = a.n();
  double b_norm = std::sqrt(dot8(b, b, n));
  std::fill(x, x + n, 0.0);
  *iterations = 0;
  *residual = 0.0;
  if (b_norm == 0.0) return OK;
  std::vector<double> r(b, b + n), z(n), p(n), ap(n);
  a.precondition(l, r.data(), z.data());
  p = z;
  double rz = dot8(r.data(), z.data(), n);
  for (i64 it = 0; it < maxit; ++it) {
    a.multiply(p.data(), ap.data());
    double curvature = dot8(p.data(), ap.data(), n);
    if (std::isnan(curvature) || curvature <= 0.0) { *iterations = it + 1; return SINGULAR; }
    double alpha = rz / curvature;
    for (i64 k = 0; k < n; ++k) {
      x[k] += alpha * p[k];
      r[k] -= alpha * ap[k];
    }
    double rn = std::sqrt(dot8(r.data(), r.data(), n));
    *iterations = it + 1;
    *residual = rn / b_norm;
    if (rn <= tol * b_norm) return OK;
    a.precondition(l, r.data(), z.data());
    double rz_next = dot8(r.data(), z.data(), n);
    double beta = rz_next / rz;
    rz = rz_next;
    for (i64 k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
  }
  return NOT_CONVERGED;
}

// ---------------------------------------------------------------- solve.rs: skyline
struct Skyline {
  std::vector<i64> first, start;
  std::vector<double> data;
  i64 n() const { return (i64)first.size(); }
  // solve.rs:19-31 (SkylineMatrix::zeros)
  void zeros(std::vector<i64> f) {
    first = std::move(f);
    start.assign(first.size() + 1, 0);
    for (size_t i = 0; i < first.size(); ++i) start[i + 1] = start[i] + (i64)i + 1 - first[i];
    data.assign(start[first.size()], 0.0);
  }
  // solve.rs:47-54
  double get(i64 i, i64 j) const {
    if (j > i) std::swap(i, j);
    return j < first[i] ? 0.0 : data[start[i] + j - first[i]];
  }
  // solve.rs:57-60
  void add(i64 i, i64 j, double v) {
    if (j > i) std::swap(i, j);
    data[start[i] + j - first[i]] += v;
  }
  // solve.rs:70-94 (factorize)
  int factorize() {
    const i64 nn = n();
    double max_diag = 0.0;
    for (i64 i = 0; i < nn; ++i) max_diag = std::fmax(max_diag, get(i, i));
    double tol = 1e-12 * max_diag;
    for (i64 i = 0; i < nn; ++i) {
      i64 fi = first[i];
      for (i64 j = fi; j <= i; ++j) {
        i64 k0 = std::max(fi, first[j]);
        i64 ri = start[i] - fi, rj = start[j] - first[j];
        double s = dot8(&data[ri + k0], &data[rj + k0], j - k0);
        double a = data[ri + j] - s;
        if (i == j) {
          if (std::isnan(a) || a <= tol) return SINGULAR;
          data[ri + j] = std::sqrt(a);
        } else {
          data[ri + j] = a / data[rj + j];
        }
      }
    }
    return OK;
  }
  // solve.rs:118-138 (Cholesky::solve)
  void solve(const double* b, double* x) const {
    const i64 nn = n();
    std::copy(b, b + nn, x);
    for (i64 i = 0; i < nn; ++i) {
      const double* row = &data[start[i]];
      i64 fi = first[i];
      double diag = row[i - fi];
      double s = 0.0;
      for (i64 k = fi; k < i; ++k) s += row[k - fi] * x[k];
      x[i] = (x[i] - s) / diag;
    }
    for (i64 i = nn - 1; i >= 0; --i) {
      const double* row = &data[start[i]];
      i64 fi = first[i];
      x[i] /= row[i - fi];
      for (i64 k = fi; k < i; ++k) x[k] -= row[k - fi] * x[i];
    }
  }
};

// ---------------------------------------------------------------- model.rs: FreeSystem
struct Mesh {
  i64 n_nodes, n_el;
  int nen;
  const double* xyz;   // n_nodes x 3
  const i64* conn;     // n_el x nen
  const double* young; // n_el
  const double* nu;    // n_el
  void coords(i64 e, double (*x)[3]) const {
    for (int a = 0; a < nen; ++a)
      for (int k = 0; k < 3; ++k) x[a][k] = xyz[3 * conn[e * nen + a] + k];
  }
  // assembly.rs:64-79 -> SolidKernel::stiffness (:241-248)
  void stiffness(i64 e, double* ke) const {
    double x[10][3], d[6][6];
    coords(e, x);
    solid_d(young[e], nu[e], d);
    solid_stiffness(nen, x, d, ke);
  }
  int dpn = 3;  // DOFs per node: 3 (displacements) or 1 (temperature, model.rs:1074-1078)
  // assembly.rs:20-32 (element_dofs, Tet branch), or the element's nodes in thermal analysis
  i64 dof(i64 e, int local) const {
    return dpn == 3 ? 3 * conn[e * nen + local / 3] + local % 3 : conn[e * nen + local];
  }
};

struct System {
  Mesh mesh;
  bool iterative = true;
  i64 ndof = 0, count = 0;
  std::vector<i64> free_of;  // global dof -> free index or -1
  std::vector<i64> order, rank;
  Csr csr;
  Skyline sky;
  std::vector<double> l;
  std::vector<double> coupling;
  std::vector<double> imposed;
  std::vector<char> fixed;
  const double* tangents = nullptr;  // n_el x m x m row-major element matrices, else stiffness
  int shift_index = -1;
  bool factored = false;
  double t_pattern = 0, t_assembly = 0, t_factor = 0, t_solve = 0;
};

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// model.rs:1442-1499 (FreeSystem::new) up to the factorisation
int system_build(System& s) {
  const Mesh& m = s.mesh;
  const int ml = m.dpn * m.nen;
  double t0 = now();
  s.free_of.assign(s.ndof, -1);
  s.count = 0;
  for (i64 g = 0; g < s.ndof; ++g)
    if (!s.fixed[g]) s.free_of[g] = s.count++;
  std::vector<i64> eoff(m.n_el + 1, 0), eit;
  eit.reserve(m.n_el * ml);
  for (i64 e = 0; e < m.n_el; ++e) {
    for (int a = 0; a < ml; ++a) {
      i64 f = s.free_of[m.dof(e, a)];
      if (f >= 0) eit.push_back(f);
    }
    eoff[e + 1] = (i64)eit.size();
  }
  Graph graph = graph_from_elements(s.count, m.n_el, eoff.data(), eit.data());
  s.order = rcm(graph);
  s.rank.assign(s.count, 0);
  for (i64 r = 0; r < s.count; ++r) s.rank[s.order[r]] = r;
  if (s.iterative) {
    s.csr = Csr::zeros(graph, s.order, s.rank);
  } else {
    std::vector<i64> first(s.count);
    for (i64 i = 0; i < s.count; ++i) first[i] = i;
    for (i64 a = 0; a < s.count; ++a)
      for (i64 p = graph.offsets[a]; p < graph.offsets[a + 1]; ++p)
        first[s.rank[a]] = std::min(first[s.rank[a]], s.rank[graph.targets[p]]);
    s.sky.zeros(std::move(first));
  }
  double t1 = now();
  s.t_pattern = t1 - t0;
  s.coupling.assign(s.count, 0.0);
  std::vector<double> ke(ml * ml);
  for (i64 e = 0; e < m.n_el; ++e) {
    if (s.tangents) std::copy(s.tangents + e * ml * ml, s.tangents + (e + 1) * ml * ml, ke.begin());
    else m.stiffness(e, ke.data());
    for (int li = 0; li < ml; ++li) {
      i64 a = s.free_of[m.dof(e, li)];
      if (a < 0) continue;
      for (int lj = 0; lj < ml; ++lj) {
        i64 gj = m.dof(e, lj);
        i64 b = s.free_of[gj];
        double k = ke[li * ml + lj];
        if (b >= 0) {
          if (s.rank[b] <= s.rank[a]) {
            if (s.iterative) {
              if (!s.csr.add(s.rank[a], s.rank[b], k)) return OUT_OF_PATTERN;
            } else {
              s.sky.add(s.rank[a], s.rank[b], k);
            }
          }
        } else {
          s.coupling[s.rank[a]] += k * s.imposed[gj];
        }
      }
    }
  }
  s.t_assembly = now() - t1;
  return OK;
}

// model.rs:1496 (kff.factorize) -> Pcg::new or SkylineMatrix::factorize
int system_factorize(System& s) {
  double t0 = now();
  int st = s.iterative ? pcg_new(s.csr, s.l, &s.shift_index) : s.sky.factorize();
  s.t_factor = now() - t0;
  s.factored = (st == OK);
  return st;
}

// model.rs:1502-1516 (FreeSystem::solve)
int system_solve(System& s, const double* f, double* u, double tol, i64 maxit, i64* iters,
                 double* resid) {
  double t0 = now();
  std::vector<double> rhs(s.count), x(s.count);
  for (i64 k = 0; k < s.count; ++k) rhs[k] = -s.coupling[k];
  for (i64 g = 0; g < s.ndof; ++g)
    if (s.free_of[g] >= 0) rhs[s.rank[s.free_of[g]]] += f[g];
  int st = OK;
  *iters = 0;
  *resid = 0.0;
  if (s.iterative) st = pcg_solve(s.csr, s.l, rhs.data(), x.data(), tol, maxit, iters, resid);
  else s.sky.solve(rhs.data(), x.data());
  for (i64 g = 0; g < s.ndof; ++g) u[g] = s.imposed[g];
  for (i64 g = 0; g < s.ndof; ++g)
    if (s.free_of[g] >= 0) u[g] = x[s.rank[s.free_of[g]]];
  s.t_solve = now() - t0;
  return st;
}

// ---------------------------------------------------------------- model.rs: quasi-static
// model.rs:1220-1224
const double NEWTON_TOLERANCE = 1e-8;
const int NEWTON_ITERATIONS = 25;
const int MAX_CUTS = 5;
enum QsStatus { NEWTON_NOT_CONVERGED = 9, PLASTIC_COLLAPSE = 10 };

// model.rs:181-247 (Amplitude::at): piecewise linear, constant after the last point
double amplitude_at(const std::vector<double>& ts, const std::vector<double>& vs, double t) {
  size_t after = 0;
  while (after < ts.size() && ts[after] <= t) ++after;  // partition_point(ti <= t)
  if (after >= 1 && after < ts.size()) {
    double t0 = ts[after - 1], v0 = vs[after - 1], t1 = ts[after], v1 = vs[after];
    return v0 + (v1 - v0) * (t - t0) / (t1 - t0);
  }
  if (after >= 1) return vs[after - 1];
  if (!ts.empty()) return vs[0];
  return 0.0;
}

// model.rs:1236-1238 (norm): sequential sum of squares
double norm_seq(const double* v, i64 n, const char* mask, bool want) {
  double s = 0.0;
  for (i64 k = 0; k < n; ++k)
    if (!mask || (mask[k] != 0) == want) s += v[k] * v[k];
  return std::sqrt(s);
}

// Model::solve_quasi_static for small strains on tetrahedra without contact, follower pressures
// or temperatures (model.rs:732-1022): the state of one run.
struct QuasiStatic {
  Mesh mesh;
  const double* yield0;     // n_el (infinite: elastic)
  const double* hardening;  // n_el
  // Kinematics::Large (total Lagrangian, finite_strain.rs): per element 1 = Neo-Hookean, 0 =
  // Saint-Venant-Kirchhoff; follower pressures (element, face, value), model.rs:868-903
  bool large = false;
  const uint8_t* neo = nullptr;
  i64 npres = 0;
  const i64* pres_el = nullptr;
  const i64* pres_face = nullptr;
  const double* pres_val = nullptr;
  std::vector<char> fixed;
  std::vector<double> imposed, f;
  std::vector<double> amp_t, amp_v;
  bool iterative = true;
  int ngp = 1;
  i64 ndof = 0;
  // QuasiStaticState (model.rs:1226-1234)
  std::vector<double> u;
  std::vector<PlasticState> committed;
  double scale = 0.0;
  // the last equilibrium: responses, reactions, residuals
  std::vector<double> stresses, p, reactions, residuals;
  std::vector<PlasticState> states;
  i64 newton_total = 0, cuts_total = 0, pcg_total = 0;
  double t_build = 0, t_solve = 0, t_resp = 0;

  // model.rs:1058-1067 (responses) with model.rs:1241-1250 (internal_force)
  bool responses(const std::vector<double>& uu, std::vector<double>& fint_g,
                 std::vector<double>* tangents) {
    const int nen = mesh.nen, m = 3 * nen;
    std::vector<double> fint(m), tan(m * m), out(9 * ngp), ue(m);
    stresses.assign(mesh.n_el * ngp * 9, 0.0);
    p.assign(mesh.n_el * ngp, 0.0);
    states.assign(mesh.n_el * ngp, PlasticState{});
    if (tangents) tangents->assign(mesh.n_el * m * m, 0.0);
    fint_g.assign(ndof, 0.0);
    bool yielding = false;
    for (i64 e = 0; e < mesh.n_el; ++e) {
      double x[10][3];
      mesh.coords(e, x);
      for (int k = 0; k < m; ++k) ue[k] = uu[mesh.dof(e, k)];
      if (large) {  // element_large_response (assembly.rs:347-355); states stay default
        const double E = mesh.young[e], v = mesh.nu[e];
        const double lam = E * v / ((1.0 + v) * (1.0 - 2.0 * v)), mu = E / (2.0 * (1.0 + v));
        large_strain_response(nen, x, neo && neo[e], lam, mu, ue.data(), fint.data(), tan.data(),
                              out.data());
      } else {
        yielding |= solid_response(nen, x, mesh.young[e], mesh.nu[e], yield0[e], hardening[e],
                                   ue.data(), &committed[e * ngp], fint.data(), tan.data(),
                                   out.data(), &states[e * ngp]);
      }
      std::copy(out.begin(), out.end(), stresses.begin() + e * ngp * 9);
      for (int q = 0; q < ngp; ++q) p[e * ngp + q] = states[e * ngp + q].p;
      if (tangents) std::copy(tan.begin(), tan.end(), tangents->begin() + e * m * m);
      for (int k = 0; k < m; ++k) fint_g[mesh.dof(e, k)] += fint[k];
    }
    return yielding;
  }

  // model.rs:923-1022 (equilibrium) from this->u: status; on success this->u is the equilibrium
  int equilibrium(std::vector<double>& uu, double lambda, double& sc) {
    residuals.clear();
    std::vector<double> fint, tangents, residual(ndof), external(ndof), du(ndof), rhs(ndof);
    for (;;) {
      double t0 = now();
      bool yielding = responses(uu, fint, &tangents);
      // model.rs:868-903 (follower_force): forces per unit lambda, tangent -= lambda sym(K_p)
      std::vector<double> follower(ndof, 0.0);
      if (large) {
        const int nen = mesh.nen, m = 3 * nen;
        int pos[6];
        double pf[18], pk[324];
        for (i64 k = 0; k < npres; ++k) {
          const i64 e = pres_el[k];
          const int nf = follower_pressure(nen, mesh.conn + e * nen, mesh.xyz, (int)pres_face[k],
                                           pres_val[k], uu.data(), pos, pf, pk);
          for (int a = 0; a < nf; ++a)
            for (int c = 0; c < 3; ++c) follower[mesh.dof(e, 3 * pos[a] + c)] += pf[3 * a + c];
          const int n3 = 3 * nf;
          double* te = &tangents[e * m * m];
          for (int r = 0; r < n3; ++r)
            for (int c = 0; c < n3; ++c) {
              const double sym = 0.5 * (pk[r * n3 + c] + pk[c * n3 + r]);
              te[(3 * pos[r / 3] + r % 3) * m + 3 * pos[c / 3] + c % 3] -= lambda * sym;
            }
        }
      }
      t_resp += now() - t0;
      for (i64 g = 0; g < ndof; ++g) external[g] = lambda * (f[g] + follower[g]) + 0.0;
      for (i64 g = 0; g < ndof; ++g) residual[g] = fint[g] - external[g];
      sc = std::fmax(std::fmax(sc, norm_seq(fint.data(), ndof, nullptr, false)),
                     norm_seq(external.data(), ndof, nullptr, false));
      const double free_norm = norm_seq(residual.data(), ndof, fixed.data(), false);
      const double relative = sc > 0.0 ? free_norm / sc : free_norm;
      if (!std::isfinite(relative)) return NEWTON_NOT_CONVERGED;
      residuals.push_back(relative);
      if (residuals.size() > 1 && relative <= NEWTON_TOLERANCE) {
        reactions.assign(ndof, 0.0);
        for (i64 g = 0; g < ndof; ++g)
          if (fixed[g]) reactions[g] = residual[g];
        return OK;
      }
      if ((int)residuals.size() > NEWTON_ITERATIONS) return NEWTON_NOT_CONVERGED;
      for (i64 g = 0; g < ndof; ++g) du[g] = fixed[g] ? lambda * imposed[g] - uu[g] : 0.0;
      System sys;
      sys.mesh = mesh;
      sys.iterative = iterative;
      sys.ndof = ndof;
      sys.fixed = fixed;
      sys.imposed = du;
      sys.tangents = tangents.data();
      double t1 = now();
      int st = system_build(sys);
      if (st != OK) return st;
      st = system_factorize(sys);
      t_build += now() - t1;
      if (st == SINGULAR && yielding) return PLASTIC_COLLAPSE;
      if (st != OK) return st;
      for (i64 g = 0; g < ndof; ++g) rhs[g] = -residual[g];
      i64 iters = 0;
      double resid = 0.0;
      t1 = now();
      st = system_solve(sys, rhs.data(), du.data(), TOLERANCE, MAX_ITERATIONS, &iters, &resid);
      t_solve += now() - t1;
      pcg_total += iters;
      ++newton_total;
      if (st != OK) return st;
      for (i64 g = 0; g < ndof; ++g) uu[g] += du[g];
    }
  }

  // model.rs:800-857 (advance)
  int advance(double t0, double t1, int cuts) {
    std::vector<double> trial = u;
    double sc = scale;
    int st = equilibrium(trial, amplitude_at(amp_t, amp_v, t1), sc);
    if (st == OK) {
      u = trial;
      scale = sc;
      committed = states;
      return OK;
    }
    if ((st == NEWTON_NOT_CONVERGED || st == SINGULAR || st == PLASTIC_COLLAPSE) &&
        cuts < MAX_CUTS) {
      const double middle = (t0 + t1) / 2.0;
      ++cuts_total;
      st = advance(t0, middle, cuts + 1);
      if (st != OK) return st;
      return advance(middle, t1, cuts + 1);
    }
    return st;
  }
};

// ---------------------------------------------------------------- model.rs: transient
// model.rs:1353-1369 (product): y[g] += sum_b M_e[a][b] x[dof_b] over the elements in order,
// for the rows a whose DOF is in `rows`; elements with no DOF in rows are skipped
void product(const Mesh& m, const std::vector<double>& mats, const double* x, const char* rows,
             i64 ndof, double* y) {
  const int ml = m.dpn * m.nen;
  std::fill(y, y + ndof, 0.0);
  for (i64 e = 0; e < m.n_el; ++e) {
    bool any = false;
    for (int a = 0; a < ml; ++a) any = any || rows[m.dof(e, a)];
    if (!any) continue;
    for (int a = 0; a < ml; ++a) {
      const i64 g = m.dof(e, a);
      if (!rows[g]) continue;
      double s = 0.0;
      for (int b = 0; b < ml; ++b) s += mats[(e * ml + a) * ml + b] * x[m.dof(e, b)];
      y[g] += s;
    }
  }
}

// Model::solve_step_with (model.rs:628-706) for tetrahedra: the theta method
// (C/dt + theta K) T_{n+1} = (C/dt - (1 - theta) K) T_n + F, K with the convection faces
// (thermal_matrix, model.rs:1320-1326), the left matrix factored once.
struct Transient {
  Mesh mesh;  // dpn = 1; young = conductivity per element
  const double* rho_c;
  std::vector<double> kmat, cmat, left, right;  // n_el x nen x nen
  std::vector<double> f;
  std::vector<char> fixed;
  std::vector<double> imposed;
  double t_build = 0, t_solve = 0;
  i64 pcg_total = 0;

  // model.rs:1080-1086 + 1320-1326: conduction, then each convection face of the element in
  // the order given, added into it
  void matrices(i64 nconv, const i64* conv_el, const i64* conv_face, const double* film,
                const double* sink) {
    const int nen = mesh.nen;
    kmat.assign(mesh.n_el * nen * nen, 0.0);
    cmat.assign(mesh.n_el * nen * nen, 0.0);
    double x[10][3];
    for (i64 e = 0; e < mesh.n_el; ++e) {
      mesh.coords(e, x);
      conduction(nen, x, mesh.young[e], &kmat[e * nen * nen]);
      capacity(nen, x, rho_c[e], &cmat[e * nen * nen]);
    }
    int pos[6];
    double mat[36], load[6];
    // faces grouped per element in the order given (model.rs:1081-1086)
    for (i64 c = 0; c < nconv; ++c) {
      const i64 e = conv_el[c];
      mesh.coords(e, x);
      const int nn = face_convection(nen, x, (int)conv_face[c], film[c], sink[c], pos, mat, load);
      for (int a = 0; a < nn; ++a)
        for (int b = 0; b < nn; ++b) kmat[(e * nen + pos[a]) * nen + pos[b]] += mat[a * nn + b];
    }
  }
};

}  // namespace

// ================================================================ C ABI (for ctypes)
extern "C" {

int or_version() { return 1; }

// element.rs:193 (solid_d), d row-major 6x6
void or_solid_d(double young, double nu, double* d) {
  double dd[6][6];
  solid_d(young, nu, dd);
  std::memcpy(d, dd, sizeof(dd));
}

// shape.rs:108/138: n[nen], dn[nen*3]
void or_tet_shape(int nen, double xi, double eta, double zeta, double* n, double* dn) {
  tet_shape(nen, xi, eta, zeta, n, (double(*)[3])dn);
}

// shape.rs:167 (Mapping3d::new): returns det, writes dn_dx[nen*3]
double or_mapping3d(int nen, const double* x, const double* dn, double* dn_dx) {
  return mapping3d(nen, (const double(*)[3])x, (const double(*)[3])dn, (double(*)[3])dn_dx);
}

// element.rs:223 via assembly.rs:241: coords[nen*3] -> ke[m*m] row-major
void or_solid_stiffness(int nen, const double* x, double young, double nu, double* ke) {
  double d[6][6];
  solid_d(young, nu, d);
  solid_stiffness(nen, (const double(*)[3])x, d, ke);
}

// element.rs:270: stresses at the Gauss points (9 values each)
void or_solid_stresses(int nen, const double* x, double young, double nu, const double* ue,
                       double* out) {
  double d[6][6];
  solid_d(young, nu, d);
  solid_stresses(nen, (const double(*)[3])x, d, ue, out);
}

// assembly.rs:650: returns the face node count; pos[6], forces[18]
int or_face_pressure(int nen, const double* xe, int face, double value, int* pos,
                     double* forces) {
  if (face < 0 || face > 3) return 0;
  return face_pressure(nen, (const double(*)[3])xe, face, value, pos, forces);
}

double or_dot(const double* a, const double* b, i64 n) { return dot8(a, b, n); }

// ---- Graph / RCM on explicit arrays (iterative.rs:38, solve.rs:200)
void* or_graph_new(i64 n, i64 ne, const i64* eoff, const i64* eit) {
  return new Graph(graph_from_elements(n, ne, eoff, eit));
}
void* or_graph_from_csr(i64 n, const i64* off, const i64* tgt) {
  Graph* g = new Graph;
  g->offsets.assign(off, off + n + 1);
  g->targets.assign(tgt, tgt + off[n]);
  return g;
}
i64 or_graph_ntargets(void* h) { return (i64)((Graph*)h)->targets.size(); }
void or_graph_get(void* h, i64* off, i64* tgt) {
  Graph* g = (Graph*)h;
  std::copy(g->offsets.begin(), g->offsets.end(), off);
  std::copy(g->targets.begin(), g->targets.end(), tgt);
}
void or_graph_rcm(void* h, i64* order) {
  auto o = rcm(*(Graph*)h);
  std::copy(o.begin(), o.end(), order);
}
void or_graph_free(void* h) { delete (Graph*)h; }

// ---- Csr on explicit arrays (iterative.rs:82-188) for the unit tests
void* or_csr_from_graph(void* gh, const i64* order) {
  Graph* g = (Graph*)gh;
  std::vector<i64> o(order, order + g->n()), r(g->n());
  for (i64 k = 0; k < g->n(); ++k) r[o[k]] = k;
  return new Csr(Csr::zeros(*g, o, r));
}
i64 or_csr_nnz(void* h) { return (i64)((Csr*)h)->cols.size(); }
int or_csr_add(void* h, i64 i, i64 j, double v) { return ((Csr*)h)->add(i, j, v) ? OK : OUT_OF_PATTERN; }
void or_csr_get(void* h, i64* off, i64* cols, double* vals) {
  Csr* a = (Csr*)h;
  if (off) std::copy(a->offsets.begin(), a->offsets.end(), off);
  if (cols) std::copy(a->cols.begin(), a->cols.end(), cols);
  if (vals) std::copy(a->values.begin(), a->values.end(), vals);
}
void* or_csr_from_arrays(i64 n, const i64* off, const i64* cols, const double* vals) {
  Csr* a = new Csr;
  a->offsets.assign(off, off + n + 1);
  a->cols.assign(cols, cols + off[n]);
  a->values.assign(vals, vals + off[n]);
  return a;
}
void or_csr_multiply(void* h, const double* x, double* y) { ((Csr*)h)->multiply(x, y); }
// IC(0) with one shift: returns 1 on success and writes l[nnz]
int or_csr_ic0(void* h, double shift, double* l) {
  std::vector<double> lv;
  if (!((Csr*)h)->incomplete_cholesky(shift, lv)) return 0;
  std::copy(lv.begin(), lv.end(), l);
  return 1;
}
void or_csr_precondition(void* h, const double* l, const double* r, double* z) {
  Csr* a = (Csr*)h;
  std::vector<double> lv(l, l + a->values.size());
  a->precondition(lv, r, z);
}
// Pcg::new + Pcg::solve on the Csr: status, iterations, relative residual, shift index
int or_csr_pcg(void* h, const double* b, double* x, double tol, i64 maxit, i64* iters,
               double* resid, int* shift_index) {
  Csr* a = (Csr*)h;
  std::vector<double> l;
  int st = pcg_new(*a, l, shift_index);
  if (st != OK) return st;
  return pcg_solve(*a, l, b, x, tol, maxit, iters, resid);
}
void or_csr_free(void* h) { delete (Csr*)h; }

// ---- dense skyline (solve.rs:10-139) on an explicit profile
void* or_sky_new(i64 n, const i64* first) {
  Skyline* s = new Skyline;
  s->zeros(std::vector<i64>(first, first + n));
  return s;
}
void or_sky_add(void* h, i64 i, i64 j, double v) { ((Skyline*)h)->add(i, j, v); }
i64 or_sky_stored(void* h) { return (i64)((Skyline*)h)->data.size(); }
double or_sky_get(void* h, i64 i, i64 j) { return ((Skyline*)h)->get(i, j); }
int or_sky_factorize(void* h) { return ((Skyline*)h)->factorize(); }
void or_sky_solve(void* h, const double* b, double* x) { ((Skyline*)h)->solve(b, x); }
void or_sky_free(void* h) { delete (Skyline*)h; }

// ---- the static linear path (model.rs:548-610) on a Tet4/Tet10 mesh
// fixed/imposed: per global DOF (3 per node). iterative: 1 = Csr + Pcg, 0 = skyline.
void* or_system_new(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                    const double* young, const double* nu, const uint8_t* fixed,
                    const double* imposed, int iterative, int* status) {
  System* s = new System;
  s->mesh = Mesh{n_nodes, n_el, nen, xyz, conn, young, nu};
  s->iterative = iterative != 0;
  s->ndof = 3 * n_nodes;
  s->fixed.assign(fixed, fixed + s->ndof);
  s->imposed.assign(imposed, imposed + s->ndof);
  *status = system_build(*s);
  return s;
}
i64 or_system_nfree(void* h) { return ((System*)h)->count; }
i64 or_system_nnz(void* h) {
  System* s = (System*)h;
  return s->iterative ? (i64)s->csr.cols.size() : (i64)s->sky.data.size();
}
void or_system_order(void* h, i64* order) {
  System* s = (System*)h;
  std::copy(s->order.begin(), s->order.end(), order);
}
void or_system_csr(void* h, i64* off, i64* cols, double* vals) {
  System* s = (System*)h;
  if (off) std::copy(s->csr.offsets.begin(), s->csr.offsets.end(), off);
  if (cols) std::copy(s->csr.cols.begin(), s->csr.cols.end(), cols);
  if (vals) std::copy(s->csr.values.begin(), s->csr.values.end(), vals);
}
void or_system_coupling(void* h, double* c) {
  System* s = (System*)h;
  std::copy(s->coupling.begin(), s->coupling.end(), c);
}
int or_system_factorize(void* h, int* shift_index) {
  System* s = (System*)h;
  int st = system_factorize(*s);
  *shift_index = s->shift_index;
  return st;
}
void or_system_lvalues(void* h, double* l) {
  System* s = (System*)h;
  std::copy(s->l.begin(), s->l.end(), l);
}
int or_system_solve(void* h, const double* f, double* u, double tol, i64 maxit, i64* iters,
                    double* resid) {
  System* s = (System*)h;
  if (!s->factored) return INVALID;
  return system_solve(*s, f, u, tol, maxit, iters, resid);
}
// phase wall times: pattern (graph + RCM + Csr::zeros), assembly, factor, solve
void or_system_times(void* h, double* t) {
  System* s = (System*)h;
  t[0] = s->t_pattern; t[1] = s->t_assembly; t[2] = s->t_factor; t[3] = s->t_solve;
}
void or_system_free(void* h) { delete (System*)h; }

// model.rs:715-797 (solve_quasi_static) for small strains on tetrahedra. Per element: E, nu,
// yield stress (infinite: elastic) and hardening. Per DOF: fixed, BC value, load F. Amplitude
// points (t, lambda). Outputs for increments 0..n (arrays sized (n+1) x ...): u, reactions,
// Gauss-point stresses (9 each) and p; residuals per increment n x 26 with their counts;
// counters[0..3] = Newton iterations, increment cuts, PCG iterations (all attempts); on a typed
// failure *fail_increment is the step increment that failed.
int or_quasi_static(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                    const double* young, const double* nu, const double* yield0,
                    const double* hardening, const uint8_t* fixed, const double* imposed,
                    const double* f, double time_period, i64 increments, i64 npts,
                    const double* amp_t, const double* amp_v, int iterative, double* u_out,
                    double* r_out, double* s_out, double* p_out, double* res_out, i64* res_len,
                    i64* counters, i64* fail_increment, double* times) {
  QuasiStatic qs;
  qs.mesh = Mesh{n_nodes, n_el, nen, xyz, conn, young, nu};
  qs.yield0 = yield0;
  qs.hardening = hardening;
  qs.ndof = 3 * n_nodes;
  qs.ngp = nen == 4 ? 1 : 4;
  qs.iterative = iterative != 0;
  qs.fixed.assign(fixed, fixed + qs.ndof);
  qs.imposed.assign(imposed, imposed + qs.ndof);
  qs.f.assign(f, f + qs.ndof);
  qs.amp_t.assign(amp_t, amp_t + npts);
  qs.amp_v.assign(amp_v, amp_v + npts);
  qs.u.assign(qs.ndof, 0.0);
  qs.committed.assign(n_el * qs.ngp, PlasticState{});
  const i64 ns = n_el * qs.ngp;
  auto record = [&](i64 k) {
    std::copy(qs.u.begin(), qs.u.end(), u_out + k * qs.ndof);
    if (r_out) std::copy(qs.reactions.begin(), qs.reactions.end(), r_out + k * qs.ndof);
    if (s_out) std::copy(qs.stresses.begin(), qs.stresses.end(), s_out + k * ns * 9);
    if (p_out) std::copy(qs.p.begin(), qs.p.end(), p_out + k * ns);
  };
  std::vector<double> fint;
  qs.responses(qs.u, fint, nullptr);  // model.rs:767 (at_rest)
  qs.reactions.assign(qs.ndof, 0.0);
  record(0);
  *fail_increment = -1;
  const double dt = time_period / (double)increments;  // Step::time_step (model.rs:171-173)
  int st = OK;
  for (i64 inc = 1; inc <= increments; ++inc) {
    const double time = (double)inc * dt;
    const double start = time - dt;
    st = qs.advance(start, time, 0);
    if (st != OK) { *fail_increment = inc; break; }
    record(inc);
    res_len[inc - 1] = (i64)qs.residuals.size();
    std::copy(qs.residuals.begin(), qs.residuals.end(), res_out + (inc - 1) * (NEWTON_ITERATIONS + 1));
  }
  counters[0] = qs.newton_total;
  counters[1] = qs.cuts_total;
  counters[2] = qs.pcg_total;
  if (times) { times[0] = qs.t_resp; times[1] = qs.t_build; times[2] = qs.t_solve; }
  return st;
}

// model.rs:1089-1127 (assemble, thermal part): F = nodal heat loads (already in f), then the
// heat sources (element, q) in order, then the convection loads h T_inf int N, faces grouped
// per element in element order (model.rs:1122-1126).
void or_thermal_loads(i64 n_nodes, const double* xyz, int nen, const i64* conn, i64 nsrc,
                      const i64* src_el, const double* src_q, i64 nconv, const i64* conv_el,
                      const i64* conv_face, const double* film, const double* sink, double* f) {
  Mesh m{n_nodes, 0, nen, xyz, conn, nullptr, nullptr};
  double x[10][3], fe[10];
  for (i64 k = 0; k < nsrc; ++k) {
    m.coords(src_el[k], x);
    nodal_source(nen, x, src_q[k], fe);
    for (int a = 0; a < nen; ++a) f[conn[src_el[k] * nen + a]] += fe[a];
  }
  std::vector<i64> order(nconv);
  for (i64 c = 0; c < nconv; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(),
                   [&](i64 a, i64 b) { return conv_el[a] < conv_el[b]; });
  int pos[6];
  double mat[36], load[6];
  for (i64 c : order) {
    m.coords(conv_el[c], x);
    const int nn = face_convection(nen, x, (int)conv_face[c], film[c], sink[c], pos, mat, load);
    for (int a = 0; a < nn; ++a) f[conn[conv_el[c] * nen + pos[a]]] += load[a];
  }
}

// Model::solve_step (model.rs:620-706) for tetrahedra. Per element: conductivity k and rho*c;
// convection faces (element, face, film, sink); per node: fixed, BC value, F (thermal loads),
// initial temperature. Outputs for increments 0..n: temperatures, reactions ((n+1) x n_nodes),
// heat fluxes ((n+1) x n_el x ngp x 6: x, y, z, qx, qy, qz). times[0..1]: build+factor, steps.
int or_transient(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                 const double* k, const double* rho_c, i64 nconv, const i64* conv_el,
                 const i64* conv_face, const double* film, const double* sink,
                 const uint8_t* fixed, const double* imposed, const double* f,
                 const double* initial, double time_period, i64 increments, double theta,
                 int iterative, double* t_out, double* r_out, double* q_out, i64* pcg_iters,
                 double* times) {
  Transient tr;
  tr.mesh = Mesh{n_nodes, n_el, nen, xyz, conn, k, nullptr};
  tr.mesh.dpn = 1;
  tr.rho_c = rho_c;
  tr.matrices(nconv, conv_el, conv_face, film, sink);
  const double dt = time_period / (double)increments;  // Step::time_step
  const i64 mm = (i64)nen * nen;
  tr.left.resize(n_el * mm);
  tr.right.resize(n_el * mm);
  for (i64 q = 0; q < n_el * mm; ++q) {  // model.rs:652-663 (combine)
    tr.left[q] = tr.cmat[q] / dt + theta * tr.kmat[q];
    tr.right[q] = tr.cmat[q] / dt + (theta - 1.0) * tr.kmat[q];
  }
  std::vector<double> t(initial, initial + n_nodes);
  for (i64 g = 0; g < n_nodes; ++g)
    if (fixed[g]) t[g] = imposed[g];
  System sys;
  sys.mesh = tr.mesh;
  sys.iterative = iterative != 0;
  sys.ndof = n_nodes;
  sys.fixed.assign(fixed, fixed + n_nodes);
  sys.imposed = t;
  sys.tangents = tr.left.data();
  double t0 = now();
  int st = system_build(sys);
  if (st == OK) st = system_factorize(sys);
  if (times) times[0] = now() - t0;
  if (st != OK) return st;
  const int ngp = nen == 4 ? 1 : 4;
  auto fluxes = [&](i64 inc) {
    if (!q_out) return;
    double x[10][3], te[10];
    for (i64 e = 0; e < n_el; ++e) {
      tr.mesh.coords(e, x);
      for (int a = 0; a < nen; ++a) te[a] = t[conn[e * nen + a]];
      heat_fluxes(nen, x, k[e], te, q_out + ((inc * n_el + e) * ngp) * 6);
    }
  };
  std::copy(t.begin(), t.end(), t_out);
  if (r_out) std::fill(r_out, r_out + n_nodes, 0.0);
  fluxes(0);
  std::vector<char> all(n_nodes, 1);
  std::vector<double> carried(n_nodes), rhs(n_nodes), react(n_nodes), u(n_nodes);
  t0 = now();
  *pcg_iters = 0;
  for (i64 inc = 1; inc <= increments; ++inc) {
    product(tr.mesh, tr.right, t.data(), all.data(), n_nodes, carried.data());
    for (i64 g = 0; g < n_nodes; ++g) rhs[g] = carried[g] + f[g];
    i64 iters = 0;
    double resid = 0.0;
    st = system_solve(sys, rhs.data(), u.data(), TOLERANCE, MAX_ITERATIONS, &iters, &resid);
    *pcg_iters += iters;
    if (st != OK) return st;
    for (i64 g = 0; g < n_nodes; ++g)
      if (!fixed[g]) t[g] = u[g];
    product(tr.mesh, tr.left, t.data(), sys.fixed.data(), n_nodes, react.data());
    for (i64 g = 0; g < n_nodes; ++g)
      if (fixed[g]) react[g] -= rhs[g];
    std::copy(t.begin(), t.end(), t_out + inc * n_nodes);
    if (r_out) std::copy(react.begin(), react.end(), r_out + inc * n_nodes);
    fluxes(inc);
  }
  if (times) times[1] = now() - t0;
  return OK;
}

// Model::solve for a thermal model (model.rs:548-610, Analysis::Thermal): K T = F with K the
// conduction + convection matrices, the BC temperatures at the fixed nodes; reactions K T - F
// on the fixed nodes (model.rs:586-593).
int or_thermal_steady(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                      const double* k, i64 nconv, const i64* conv_el, const i64* conv_face,
                      const double* film, const double* sink, const uint8_t* fixed,
                      const double* imposed, const double* f, int iterative, double* t,
                      double* r) {
  Transient tr;
  tr.mesh = Mesh{n_nodes, n_el, nen, xyz, conn, k, nullptr};
  tr.mesh.dpn = 1;
  std::vector<double> ones(n_el, 1.0);
  tr.rho_c = ones.data();
  tr.matrices(nconv, conv_el, conv_face, film, sink);
  System sys;
  sys.mesh = tr.mesh;
  sys.iterative = iterative != 0;
  sys.ndof = n_nodes;
  sys.fixed.assign(fixed, fixed + n_nodes);
  sys.imposed.assign(imposed, imposed + n_nodes);
  sys.tangents = tr.kmat.data();
  int st = system_build(sys);
  if (st == OK) st = system_factorize(sys);
  if (st != OK) return st;
  i64 iters = 0;
  double resid = 0.0;
  st = system_solve(sys, f, t, TOLERANCE, MAX_ITERATIONS, &iters, &resid);
  if (st != OK) return st;
  product(tr.mesh, tr.kmat, t, sys.fixed.data(), n_nodes, r);
  for (i64 g = 0; g < n_nodes; ++g)
    if (fixed[g]) r[g] -= f[g];
  return OK;
}

// element matrices for the unit tests: conduction k (nen x nen), capacity, nodal source
void or_thermal_element(int nen, const double* x, double k, double rho_c, double q, double* ke,
                        double* ce, double* fe) {
  const double (*xe)[3] = (const double (*)[3])x;
  conduction(nen, xe, k, ke);
  capacity(nen, xe, rho_c, ce);
  nodal_source(nen, xe, q, fe);
}

// Model::solve_quasi_static with Kinematics::Large (model.rs:732-1022 with element_large_response
// and follower_force) on tetrahedra: per element E, nu and the law (neo[e] = 1: Neo-Hookean, else
// Saint-Venant-Kirchhoff); follower pressures (element, face, value); outputs as or_quasi_static
// (stresses are Cauchy stresses at the current Gauss-point positions; p stays 0).
int or_large_strain(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                    const double* young, const double* nu, const uint8_t* neo, i64 npres,
                    const i64* pres_el, const i64* pres_face, const double* pres_val,
                    const uint8_t* fixed, const double* imposed, const double* f,
                    double time_period, i64 increments, i64 npts, const double* amp_t,
                    const double* amp_v, int iterative, double* u_out, double* r_out,
                    double* s_out, double* res_out, i64* res_len, i64* counters,
                    i64* fail_increment) {
  std::vector<double> inf(n_el, INFINITY), zero(n_el, 0.0);
  QuasiStatic qs;
  qs.mesh = Mesh{n_nodes, n_el, nen, xyz, conn, young, nu};
  qs.yield0 = inf.data();
  qs.hardening = zero.data();
  qs.large = true;
  qs.neo = neo;
  qs.npres = npres;
  qs.pres_el = pres_el;
  qs.pres_face = pres_face;
  qs.pres_val = pres_val;
  qs.ndof = 3 * n_nodes;
  qs.ngp = nen == 4 ? 1 : 4;
  qs.iterative = iterative != 0;
  qs.fixed.assign(fixed, fixed + qs.ndof);
  qs.imposed.assign(imposed, imposed + qs.ndof);
  qs.f.assign(f, f + qs.ndof);
  qs.amp_t.assign(amp_t, amp_t + npts);
  qs.amp_v.assign(amp_v, amp_v + npts);
  qs.u.assign(qs.ndof, 0.0);
  qs.committed.assign(n_el * qs.ngp, PlasticState{});
  const i64 ns = n_el * qs.ngp;
  auto record = [&](i64 k) {
    std::copy(qs.u.begin(), qs.u.end(), u_out + k * qs.ndof);
    if (r_out) std::copy(qs.reactions.begin(), qs.reactions.end(), r_out + k * qs.ndof);
    if (s_out) std::copy(qs.stresses.begin(), qs.stresses.end(), s_out + k * ns * 9);
  };
  std::vector<double> fint;
  qs.responses(qs.u, fint, nullptr);
  qs.reactions.assign(qs.ndof, 0.0);
  record(0);
  *fail_increment = -1;
  const double dt = time_period / (double)increments;
  int st = OK;
  for (i64 inc = 1; inc <= increments; ++inc) {
    const double time = (double)inc * dt;
    st = qs.advance(time - dt, time, 0);
    if (st != OK) { *fail_increment = inc; break; }
    record(inc);
    res_len[inc - 1] = (i64)qs.residuals.size();
    std::copy(qs.residuals.begin(), qs.residuals.end(), res_out + (inc - 1) * (NEWTON_ITERATIONS + 1));
  }
  counters[0] = qs.newton_total;
  counters[1] = qs.cuts_total;
  counters[2] = qs.pcg_total;
  return st;
}

// finite_strain.rs:119-214 for one element (unit tests): fint[m], tan[m*m], out[ngp*9]
void or_large_response(int nen, const double* x, int neo, double lambda, double mu,
                       const double* ue, double* fint, double* tan, double* out) {
  large_strain_response(nen, (const double (*)[3])x, neo != 0, lambda, mu, ue, fint, tan, out);
}

// assembly.rs:732-862 for one face (unit tests): returns m; pos[6], forces[3m], stiffness[9m^2]
int or_follower_pressure(int nen, const i64* enodes, const double* xyz, int face, double value,
                         const double* u, int* pos, double* forces, double* stiffness) {
  return follower_pressure(nen, enodes, xyz, face, value, u, pos, forces, stiffness);
}

// plasticity.rs:45-112 for the unit tests: stress[6], state[7] (ep, p), tangent[36]; returns yielding
int or_von_mises(double young, double poisson, double yield0, double hardening,
                 const double* strain, const double* prev, double* stress, double* state,
                 double* tangent) {
  PlasticState a, b;
  std::copy(prev, prev + 6, a.ep);
  a.p = prev[6];
  double c[6][6];
  bool y = von_mises_update(young, poisson, yield0, hardening, strain, a, stress, b, c);
  std::copy(b.ep, b.ep + 6, state);
  state[6] = b.p;
  std::memcpy(tangent, c, sizeof(c));
  return y ? 1 : 0;
}

// element.rs:395-452 for one element: fint[m], tangent[m*m] (row-major tangent[j][i]),
// out[ngp*9], states in/out [ngp*7]; returns yielding
int or_solid_response(int nen, const double* x, double young, double poisson, double yield0,
                      double hardening, const double* ue, const double* prev, double* fint,
                      double* tan, double* out, double* states) {
  int ngp = nen == 4 ? 1 : 4;
  PlasticState a[4], b[4];
  for (int q = 0; q < ngp; ++q) {
    std::copy(prev + 7 * q, prev + 7 * q + 6, a[q].ep);
    a[q].p = prev[7 * q + 6];
  }
  bool y = solid_response(nen, (const double (*)[3])x, young, poisson, yield0, hardening, ue, a,
                          fint, tan, out, b);
  for (int q = 0; q < ngp; ++q) {
    std::copy(b[q].ep, b[q].ep + 6, states + 7 * q);
    states[7 * q + 6] = b[q].p;
  }
  return y ? 1 : 0;
}

// The multi-core CPU baseline (timing only, not a parity reference): the same IC(0)-PCG on the
// factored System, with OpenMP over the host cores. SpMV row-parallel on the full symmetric
// matrix, forward and backward triangular solves level-scheduled (rows of one dependency level in
// parallel, each row summed as the reference sums it), dot products and vector updates parallel.
// Runs `iters` PCG iterations on the right-hand side of f (no stop test) and returns the wall
// time per iteration (and, in *rrel, |r|/|b| at the end); threads <= 0 uses the OpenMP default.
static double pcg_parallel_run(void* h, const double* f, i64 iters, int threads, double tol,
                               double* x_out, i64* iters_done, double* rrel) {
  System* s = (System*)h;
  if (!s->factored || !s->iterative) return -1.0;
  const Csr& a = s->csr;
  const i64 n = a.n();
  const std::vector<double>& l = s->l;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  // full symmetric rows for the SpMV
  std::vector<i64> foff(n + 1, 0);
  for (i64 i = 0; i < n; ++i)
    for (i64 p = a.offsets[i]; p < a.offsets[i + 1]; ++p) {
      foff[i + 1] += 1;
      if (a.cols[p] != i) foff[a.cols[p] + 1] += 1;
    }
  for (i64 i = 0; i < n; ++i) foff[i + 1] += foff[i];
  std::vector<i64> fcol(foff[n]), cur(foff.begin(), foff.end() - 1);
  std::vector<double> fval(foff[n]);
  // upper rows (L^T) for the backward solve: row j holds l_ij for i > j
  std::vector<i64> uoff(n + 1, 0);
  for (i64 i = 0; i < n; ++i)
    for (i64 p = a.offsets[i]; p < a.offsets[i + 1] - 1; ++p) uoff[a.cols[p] + 1] += 1;
  for (i64 i = 0; i < n; ++i) uoff[i + 1] += uoff[i];
  std::vector<i64> ucol(uoff[n]), ucur(uoff.begin(), uoff.end() - 1);
  std::vector<double> uval(uoff[n]);
  for (i64 i = 0; i < n; ++i)
    for (i64 p = a.offsets[i]; p < a.offsets[i + 1]; ++p) {
      const i64 j = a.cols[p];
      fcol[cur[i]] = j;
      fval[cur[i]++] = a.values[p];
      if (j != i) {
        fcol[cur[j]] = i;
        fval[cur[j]++] = a.values[p];
        ucol[ucur[j]] = i;
        uval[ucur[j]++] = l[p];
      }
    }
  // dependency levels of the forward (L) and backward (L^T) solves
  std::vector<i64> lev_f(n, 0), lev_b(n, 0);
  i64 depth_f = 0, depth_b = 0;
  for (i64 i = 0; i < n; ++i) {
    i64 d = 0;
    for (i64 p = a.offsets[i]; p < a.offsets[i + 1] - 1; ++p) d = std::max(d, lev_f[a.cols[p]] + 1);
    lev_f[i] = d;
    depth_f = std::max(depth_f, d + 1);
  }
  for (i64 i = n - 1; i >= 0; --i) {
    i64 d = 0;
    for (i64 p = uoff[i]; p < uoff[i + 1]; ++p) d = std::max(d, lev_b[ucol[p]] + 1);
    lev_b[i] = d;
    depth_b = std::max(depth_b, d + 1);
  }
  auto bucket = [&](const std::vector<i64>& lev, i64 depth, std::vector<i64>& off,
                    std::vector<i64>& rows) {
    off.assign(depth + 1, 0);
    for (i64 i = 0; i < n; ++i) off[lev[i] + 1] += 1;
    for (i64 d = 0; d < depth; ++d) off[d + 1] += off[d];
    rows.resize(n);
    std::vector<i64> c(off.begin(), off.end() - 1);
    for (i64 i = 0; i < n; ++i) rows[c[lev[i]]++] = i;
  };
  std::vector<i64> fo, fr, bo, br;
  bucket(lev_f, depth_f, fo, fr);
  bucket(lev_b, depth_b, bo, br);
  std::vector<double> b(n), x(n, 0.0), r(n), z(n), y(n), pv(n), ap(n);
  for (i64 k = 0; k < n; ++k) b[k] = -s->coupling[k];
  for (i64 g = 0; g < s->ndof; ++g)
    if (s->free_of[g] >= 0) b[s->rank[s->free_of[g]]] += f[g];
  double rz = 0.0, pap = 0.0, rz_next = 0.0, rr = 0.0, bb = 0.0;
  for (i64 k = 0; k < n; ++k) bb += b[k] * b[k];
  const double stop = tol > 0.0 ? tol * std::sqrt(bb) : -1.0;
  i64 done = 0;
  bool stopped = bb == 0.0 && tol > 0.0;  // b = 0: x = 0 (iterative.rs:211-216)
  const double t0 = now();
#pragma omp parallel
  {
    auto precondition = [&]() {
      for (i64 d = 0; d < depth_f; ++d) {
#pragma omp for schedule(static)
        for (i64 q = fo[d]; q < fo[d + 1]; ++q) {
          const i64 i = fr[q];
          double sum = 0.0;
          const i64 e = a.offsets[i + 1] - 1;
          for (i64 p = a.offsets[i]; p < e; ++p) sum += l[p] * y[a.cols[p]];
          y[i] = (r[i] - sum) / l[e];
        }
      }
      for (i64 d = 0; d < depth_b; ++d) {
#pragma omp for schedule(static)
        for (i64 q = bo[d]; q < bo[d + 1]; ++q) {
          const i64 i = br[q];
          double v = y[i];
          for (i64 p = uoff[i + 1] - 1; p >= uoff[i]; --p) v -= uval[p] * z[ucol[p]];
          z[i] = v / l[a.offsets[i + 1] - 1];
        }
      }
    };
#pragma omp for schedule(static)
    for (i64 k = 0; k < n; ++k) r[k] = b[k];
    precondition();
#pragma omp for schedule(static) reduction(+ : rz)
    for (i64 k = 0; k < n; ++k) {
      pv[k] = z[k];
      rz += r[k] * z[k];
    }
    for (i64 it = 0; it < iters && !stopped; ++it) {
#pragma omp single
      {
        pap = 0.0;
        rz_next = 0.0;
        rr = 0.0;
      }
#pragma omp for schedule(static) reduction(+ : pap)
      for (i64 i = 0; i < n; ++i) {
        double sum = 0.0;
        for (i64 p = foff[i]; p < foff[i + 1]; ++p) sum += fval[p] * pv[fcol[p]];
        ap[i] = sum;
        pap += pv[i] * sum;
      }
      const double alpha = rz / pap;
#pragma omp for schedule(static) reduction(+ : rr)
      for (i64 k = 0; k < n; ++k) {
        x[k] += alpha * pv[k];
        r[k] -= alpha * ap[k];
        rr += r[k] * r[k];
      }
#pragma omp single
      {
        done = it + 1;
        if (std::sqrt(rr) <= stop) stopped = true;  // iterative.rs:233-235
      }
      if (stopped) break;
      precondition();
#pragma omp for schedule(static) reduction(+ : rz_next)
      for (i64 k = 0; k < n; ++k) rz_next += r[k] * z[k];
      const double beta = rz_next / rz;
#pragma omp for schedule(static)
      for (i64 k = 0; k < n; ++k) pv[k] = z[k] + beta * pv[k];
#pragma omp single
      rz = rz_next;
    }
  }
  const double per_iter = (now() - t0) / (double)std::max<i64>(done, 1);
  if (rrel) {  // |r| / |b| after the iterations, to check the variant converges as the serial one
    double r2 = 0.0;
    for (i64 k = 0; k < n; ++k) r2 += r[k] * r[k];
    *rrel = bb > 0.0 ? std::sqrt(r2 / bb) : 0.0;
  }
  if (iters_done) *iters_done = done;
  if (x_out) std::copy(x.begin(), x.end(), x_out);
  return per_iter;
}

double or_system_pcg_parallel(void* h, const double* f, i64 iters, int threads, double* rrel) {
  return pcg_parallel_run(h, f, iters, threads, 0.0, nullptr, nullptr, rrel);
}

// The same OpenMP PCG run to the reference's stop test |r| <= tol |b| (iterative.rs:210-248)
// and scattered to u over all DOFs with the imposed values (model.rs:1502-1516): a full-size
// reference solution for the device (its dot products and sums over rows run in another
// order than the serial restatement's, so it agrees with it to rounding, not bit for bit).
int or_system_solve_parallel(void* h, const double* f, double* u, double tol, i64 maxit,
                             int threads, i64* iters, double* rrel) {
  System* s = (System*)h;
  if (!s->factored || !s->iterative) return INVALID;
  std::vector<double> x(s->count);
  i64 done = 0;
  double res = 0.0;
  pcg_parallel_run(h, f, maxit, threads, tol, x.data(), &done, &res);
  for (i64 g = 0; g < s->ndof; ++g)
    u[g] = s->free_of[g] >= 0 ? x[s->rank[s->free_of[g]]] : s->imposed[g];
  if (iters) *iters = done;
  if (rrel) *rrel = res;
  return res <= tol ? OK : NOT_CONVERGED;
}

// OpenMP threads the parallel baseline would use (1 without OpenMP)
int or_omp_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// model.rs:586-593 + :1353-1369 (product): reactions R = K u - F on the fixed DOFs, 0 elsewhere
void or_reactions(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                  const double* young, const double* nu, const uint8_t* fixed, const double* f,
                  const double* u, double* reactions) {
  Mesh m{n_nodes, n_el, nen, xyz, conn, young, nu};
  const int ml = 3 * nen;
  const i64 ndof = 3 * n_nodes;
  std::fill(reactions, reactions + ndof, 0.0);
  std::vector<double> ke(ml * ml);
  for (i64 e = 0; e < n_el; ++e) {
    bool any = false;
    for (int a = 0; a < ml; ++a) any = any || fixed[m.dof(e, a)];
    if (!any) continue;
    m.stiffness(e, ke.data());
    for (int li = 0; li < ml; ++li) {
      i64 g = m.dof(e, li);
      if (!fixed[g]) continue;
      double s = 0.0;
      for (int lj = 0; lj < ml; ++lj) s += ke[li * ml + lj] * u[m.dof(e, lj)];
      reactions[g] += s;
    }
  }
  for (i64 g = 0; g < ndof; ++g)
    if (fixed[g]) reactions[g] -= f[g];
}

// model.rs:596-598 -> element_stresses for every element: out[n_el * ngp * 9]
void or_stresses(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                 const double* young, const double* nu, const double* u, double* out) {
  Mesh m{n_nodes, n_el, nen, xyz, conn, young, nu};
  int ngp;
  rule_of(nen, &ngp);
  const int ml = 3 * nen;
  for (i64 e = 0; e < n_el; ++e) {
    double x[10][3], d[6][6], ue[30];
    m.coords(e, x);
    solid_d(young[e], nu[e], d);
    for (int a = 0; a < ml; ++a) ue[a] = u[m.dof(e, a)];
    solid_stresses(nen, x, d, ue, out + e * ngp * 9);
  }
}

// results.rs:184-221 (smooth) with :138-166 (nodal_stresses) on tetrahedra. The Gauss-point
// values gp[e][g][3..8] are extrapolated to each element's nodes and summed per node over the
// elements in element order and, inside an element, over the Gauss points in order (s += w * v),
// then divided by the number of elements at the node (an isolated node keeps zeros). Tet4: one
// Gauss point, weight 1 (results.rs:96-97). Tet10: results.rs:113-133, the TETRA_4 point g lying
// near vertex g: a vertex takes (3 sqrt5 + 1)/4 of its own point and -(sqrt5 - 1)/4 of the
// others, an edge mid-node (sqrt5 + 1)/4 of its two end points' and -(sqrt5 - 1)/4 of the others.
// out[n][0..2] = the node's position, out[n][3..8] the six components in gp's order.
static const double TET10_VERTEX = 1.9270509831248424;
static const double TET10_MIDDLE = 0.8090169943749475;
static const double TET10_OTHER = -0.30901699437494745;

static double extrapolation_weight(int nen, int k, int g) {
  if (nen == 4) return 1.0;
  if (k < 4) return k == g ? TET10_VERTEX : TET10_OTHER;
  const int a = TET10_EDGES[k - 4][0], b = TET10_EDGES[k - 4][1];
  return (g == a || g == b) ? TET10_MIDDLE : TET10_OTHER;
}

void or_nodal_stresses(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                       const double* gp, double* out) {
  const int ngp = nen == 4 ? 1 : 4;
  std::vector<double> sum(6 * n_nodes, 0.0);
  std::vector<i64> count(n_nodes, 0);
  for (i64 e = 0; e < n_el; ++e) {
    const double* points = gp + e * ngp * 9;
    for (int k = 0; k < nen; ++k) {
      const i64 n = conn[e * nen + k];
      for (int g = 0; g < ngp; ++g) {
        const double w = extrapolation_weight(nen, k, g);
        for (int c = 0; c < 6; ++c) sum[6 * n + c] += w * points[g * 9 + 3 + c];
      }
      count[n] += 1;
    }
  }
  for (i64 n = 0; n < n_nodes; ++n) {
    for (int d = 0; d < 3; ++d) out[9 * n + d] = xyz[3 * n + d];
    const double c = (double)std::max<i64>(count[n], 1);
    for (int j = 0; j < 6; ++j) out[9 * n + 3 + j] = sum[6 * n + j] / c;
  }
}

// The FP64 CPU adjoint of the compliance C = f.u (no counterpart in the reference, which has
// no sensitivities): K_e = solid_stiffness (element.rs:223-243) is linear in E_e and the
// adjoint state of C is u, so dC/dE_e = -u_e^T (K_e / E_e) u_e, summed row by row.
void or_compliance_gradient(i64 n_nodes, const double* xyz, i64 n_el, int nen, const i64* conn,
                            const double* young, const double* nu, const double* u, double* out) {
  Mesh m{n_nodes, n_el, nen, xyz, conn, young, nu};
  const int ml = 3 * nen;
  std::vector<double> ke(ml * ml);
  for (i64 e = 0; e < n_el; ++e) {
    double x[10][3], d[6][6], ue[30];
    m.coords(e, x);
    solid_d(young[e], nu[e], d);
    solid_stiffness(nen, x, d, ke.data());
    for (int a = 0; a < ml; ++a) ue[a] = u[m.dof(e, a)];
    double s = 0.0;
    for (int i = 0; i < ml; ++i) {
      double t = 0.0;
      for (int j = 0; j < ml; ++j) t += ke[i * ml + j] * ue[j];
      s += ue[i] * t;
    }
    out[e] = -s / young[e];
  }
}

// model.rs:1070-1147 (assemble), the loads part: nodal loads already in f; adds the dead
// pressure loads (element, face, value) in order through FacePressure::add_load.
void or_add_pressures(i64 n_nodes, const double* xyz, int nen, const i64* conn, i64 np,
                      const i64* p_el, const i64* p_face, const double* p_val, double* f) {
  (void)n_nodes;
  for (i64 k = 0; k < np; ++k) {
    i64 e = p_el[k];
    double xe[10][3];
    for (int a = 0; a < nen; ++a)
      for (int c = 0; c < 3; ++c) xe[a][c] = xyz[3 * conn[e * nen + a] + c];
    int pos[6];
    double forces[18];
    int nn = face_pressure(nen, xe, (int)p_face[k], p_val[k], pos, forces);
    for (int a = 0; a < nn; ++a)
      for (int c = 0; c < 3; ++c) f[3 * conn[e * nen + pos[a]] + c] += forces[3 * a + c];
  }
}

// assembly.rs:680-710 (edge_load) with FacePressure::add_load (:638-644), model.rs:1123-1127:
// the dead line load q (force per unit length) on edge `edge` of a tetrahedron, f_a = q int N_a ds:
// the edge's nodes are TET10_EDGES[edge] (and mid-node 4 + edge on a Tet10, mesh.rs:204-211),
// its length that of the straight segment between its two vertices, LINE_2 with line2 (two
// nodes) or LINE_3 with line3 (three nodes, quadrature.rs:62-90, shape.rs:115-125); each term
// q_c * w * length * N_a is added rule point after rule point, then into f node after node.
void or_add_edge_loads(i64 n_nodes, const double* xyz, int nen, const i64* conn, i64 nl,
                       const i64* l_el, const i64* l_edge, const double* q, double* f) {
  (void)n_nodes;
  static const double LINE_2_S[2] = {0.2113248654051871, 0.7886751345948129};
  static const double LINE_2_W[2] = {0.5, 0.5};
  static const double LINE_3_S[3] = {0.1127016653792583, 0.5, 0.8872983346207417};
  static const double LINE_3_W[3] = {5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0};
  for (i64 k = 0; k < nl; ++k) {
    const i64 e = l_el[k];
    const int edge = (int)l_edge[k];
    int pos[3] = {TET10_EDGES[edge][0], TET10_EDGES[edge][1], 4 + edge};
    const int m = nen == 4 ? 2 : 3;
    const double* c0 = xyz + 3 * conn[e * nen + pos[0]];
    const double* c1 = xyz + 3 * conn[e * nen + pos[1]];
    double len2 = 0.0;
    for (int c = 0; c < 3; ++c) len2 += (c1[c] - c0[c]) * (c1[c] - c0[c]);
    const double length = std::sqrt(len2);
    double forces[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int np = m == 2 ? 2 : 3;
    for (int g = 0; g < np; ++g) {
      const double sv = m == 2 ? LINE_2_S[g] : LINE_3_S[g];
      const double w = m == 2 ? LINE_2_W[g] : LINE_3_W[g];
      double na[3];
      if (m == 2) {
        na[0] = 1.0 - sv;
        na[1] = sv;
      } else {
        na[0] = (1.0 - sv) * (1.0 - 2.0 * sv);
        na[1] = sv * (2.0 * sv - 1.0);
        na[2] = 4.0 * sv * (1.0 - sv);
      }
      for (int a = 0; a < m; ++a)
        for (int c = 0; c < 3; ++c) forces[3 * a + c] += q[3 * k + c] * w * length * na[a];
    }
    for (int a = 0; a < m; ++a)
      for (int c = 0; c < 3; ++c) f[3 * conn[e * nen + pos[a]] + c] += forces[3 * a + c];
  }
}

}  // extern "C"
