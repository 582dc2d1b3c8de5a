// Closed-form 1D reference-element operators (see basis.cpp).
#pragma once
#include <vector>

namespace bpx {
namespace basis {

void legendre(int n, double x, double& P, double& dP);
double orthonormal_legendre(int n, double x, double* deriv);
void gll(int N, std::vector<double>& x, std::vector<double>& w);
void gauss(int n, std::vector<double>& x, std::vector<double>& w);  // n points
void lagrange_at(const std::vector<double>& z, double x, std::vector<double>& l);
void dmatrix_gll(int N, const std::vector<double>& x, std::vector<double>& D);
void dmatrix(const std::vector<double>& z, std::vector<double>& D);
void interp(const std::vector<double>& zin, const std::vector<double>& zout, std::vector<double>& I);
void mass_gll(int N, const std::vector<double>& x, std::vector<double>& MM);
void inv_vandermonde(int N, const std::vector<double>& x, std::vector<double>& invV);
void modal_dmatrix(int N, std::vector<double>& Dm);

}  // namespace basis

// triangle and tetrahedron operators (basis_simplex.cpp)
namespace simplex {
// warp & blend nodes in the reference's order; idx holds each node's
// lattice indices (i, j[, k]): its barycentric weights are i/N on the
// vertex (1,-1[,-1]), j/N on (-1,1[,-1]), k/N on (-1,-1,1), the rest on
// (-1,-1[,-1])
void nodes_tri(int N, std::vector<double>& r, std::vector<double>& s, std::vector<int>& idx);
void nodes_tet(int N, std::vector<double>& r, std::vector<double>& s, std::vector<double>& t,
               std::vector<int>& idx);
// orthonormal modes at the points (row = point, column = mode) and their
// r, s (, t) derivatives
void vandermonde_tri(int N, const std::vector<double>& r, const std::vector<double>& s, std::vector<double>& V,
                     std::vector<double>& Vr, std::vector<double>& Vs);
void vandermonde_tet(int N, const std::vector<double>& r, const std::vector<double>& s, const std::vector<double>& t,
                     std::vector<double>& V, std::vector<double>& Vr, std::vector<double>& Vs,
                     std::vector<double>& Vt);
void invert(int n, std::vector<double>& A);
// MM (Np x Np), D (dim x Np x Np, D_d[n][m] = d l_m / d r_d at node n) and
// the packed stiffness S (3 or 6 x Np x Np), all row-major
void operators(int Np, int dim, const std::vector<double>& V, const std::vector<const std::vector<double>*>& Vg,
               std::vector<double>& MM, std::vector<double>& D, std::vector<double>& S);
}  // namespace simplex
}  // namespace bpx
