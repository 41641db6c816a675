// SWIPDG for -div(grad u) = f on [-1, 1]^2 with nonzero Dirichlet data, written against the C++ classes of
// hgdt.hh the way examples/adaptive_elliptic_swipdg.cc:216-251 of dune-gdt assembles its SWIPDG system
// (symmetry prefactor 1, penalty 16, weight = diffusion), with the Dirichlet data added to the right-hand
// side by boundary functionals (LocalIntersectionIntegralFunctional with the Dirichlet coupling and the
// boundary penalty).  Exact solution u = cos(pi x / 2) cos(pi y / 2) + x / 4 - y / 2 (the linear part is
// harmonic, so f = pi^2 / 2 cos(pi x / 2) cos(pi y / 2) as in the ESV2007 problem).
//
// usage: elliptic_swipdg [cells per direction = 32] [solution output file (raw float64)]
// Prints the H1-semi error and one JSON line with the results.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "hgdt.hh"

using namespace hgdt;

int main(int argc, char* argv[])
{
  try {
    const int n = argc > 1 ? std::atoi(argv[1]) : 32;
    const double diffusion = 1.;
    const double symmetry_prefactor = 1; // SIPDG
    const double penalty_parameter = 16;
    const GenericFunction force(3, [](const double* x) {
      return M_PI_2 * M_PI * std::cos(M_PI_2 * x[0]) * std::cos(M_PI_2 * x[1]);
    });
    const GenericFunction exact(
        3,
        [](const double* x) { return std::cos(M_PI_2 * x[0]) * std::cos(M_PI_2 * x[1]) + 0.25 * x[0] - 0.5 * x[1]; },
        "exact_solution",
        [](const double* x, double* g) {
          g[0] = -M_PI_2 * std::sin(M_PI_2 * x[0]) * std::cos(M_PI_2 * x[1]) + 0.25;
          g[1] = -M_PI_2 * std::cos(M_PI_2 * x[0]) * std::sin(M_PI_2 * x[1]) - 0.5;
        });

    auto grid = make_cube_grid(Geometry::quad, -1., 1., n);
    const auto& grid_view = grid.leaf_view();
    AllDirichletBoundaryInfo boundary_info;
    auto dg_space = make_discontinuous_lagrange_space(grid_view, 1);

    auto lhs_op = make_matrix_operator(dg_space, Stencil::element_and_intersection);
    lhs_op.append(LocalElementIntegralBilinearForm(LocalLaplaceIntegrand(diffusion)));
    lhs_op.append(LocalCouplingIntersectionIntegralBilinearForm(
                      LocalLaplaceIPDGIntegrands::InnerCoupling(symmetry_prefactor, diffusion, diffusion)
                      + LocalIPDGIntegrands::InnerPenalty(penalty_parameter, diffusion)),
                  {},
                  ApplyOn::InnerIntersectionsOnce());
    lhs_op.append(LocalIntersectionIntegralBilinearForm(
                      LocalIPDGIntegrands::BoundaryPenalty(penalty_parameter, diffusion)
                      + LocalLaplaceIPDGIntegrands::DirichletCoupling(symmetry_prefactor, diffusion)),
                  {},
                  ApplyOn::CustomBoundaryIntersections(boundary_info));
    auto rhs_func = make_vector_functional(dg_space);
    rhs_func.append(LocalElementIntegralFunctional(LocalProductIntegrand().with_ansatz(force)));
    // the Dirichlet data, weakly: the right-hand sides of the two boundary terms of the bilinear form
    rhs_func.append(LocalIntersectionIntegralFunctional(
                        LocalLaplaceIPDGIntegrands::DirichletCoupling(symmetry_prefactor, diffusion, exact)),
                    {},
                    ApplyOn::CustomBoundaryIntersections(boundary_info));
    rhs_func.append(LocalIntersectionIntegralFunctional(
                        LocalIPDGIntegrands::BoundaryPenalty(penalty_parameter, diffusion).with_ansatz(exact)),
                    {},
                    ApplyOn::CustomBoundaryIntersections(boundary_info));
    lhs_op.append(rhs_func);
    // third argument "matrix-free": no stored matrix, the operator is applied from the coefficients
    if (argc > 3 && std::string(argv[3]) == "matrix-free")
      lhs_op.set_format(Format::matrix_free);
    lhs_op.assemble(/*use_tbb=*/true);

    auto solution = make_discrete_function(dg_space);
    SolverOptions opts;
    opts.precision = 1e-10;
    const auto stats = lhs_op.apply_inverse(lhs_op.rhs_device(), solution.dofs(), opts);

    auto h1_prod = make_bilinear_form(grid_view, solution - exact);
    h1_prod += LocalElementIntegralBilinearForm(LocalLaplaceIntegrand(diffusion), /*over_integrate=*/4);
    const double h1 = std::sqrt(h1_prod.apply2());
    std::cout << "error in H^1 semi-norm: " << h1 << std::endl;
    std::printf("{\"dofs\": %lld, \"iterations\": %lld, \"relres\": %.17g, \"h1_semi_error\": %.17g}\n",
                (long long)dg_space.size(), (long long)stats.iterations, stats.relres, h1);
    if (argc > 2) {
      const auto u = solution.dofs().to_host();
      FILE* f = std::fopen(argv[2], "wb");
      if (!f || std::fwrite(u.data(), sizeof(double), u.size(), f) != u.size()) {
        std::cerr << "cannot write " << argv[2] << std::endl;
        return EXIT_FAILURE;
      }
      std::fclose(f);
    }
  } catch (Exception& e) {
    std::cerr << "\nhgdt reported error (status " << int(e.status()) << "): " << e.what() << std::endl;
    return EXIT_FAILURE;
  } catch (std::exception& e) {
    std::cerr << "\nstl reported error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
