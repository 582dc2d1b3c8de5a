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
}  // namespace bpx
