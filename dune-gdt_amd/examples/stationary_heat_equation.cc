// Stationary heat equation -div(grad u) = f on [-1, 1]^2 with Q1 continuous Lagrange elements, written
// against the C++ classes of hgdt.hh exactly as examples/stationary-heat-equation.cc:64-132 of dune-gdt
// is written against Dune::GDT: Laplace operator, product functional with an analytic source, Dirichlet
// constraints, solve, H1-semi and L2 errors against the exact solution u = cos(pi x / 2) cos(pi y / 2).
//
// usage: stationary_heat_equation [cells per direction = 128] [solution output file (raw float64)] [grid]
// grid: "cube" (default: quadrilaterals) or "mixed" (the cells of the left half quadrilaterals, those of
// the right half split into two triangles, built element by element as GridFactory::insertElement does).
// Prints the errors and one JSON line with the results.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "hgdt.hh"

using namespace hgdt;

// [-1, 1]^2 in n x n cells: columns left of n / 2 quadrilaterals, the others two triangles each
// (v0, v0 + e_x, v0 + e_x + e_y) and (v0, v0 + e_y, v0 + e_x + e_y); vertices lexicographic, x fastest
static Grid make_mixed_strip_grid(int n)
{
  std::vector<double> coords;
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i)
      coords.push_back(-1. + 2. * i / n), coords.push_back(-1. + 2. * j / n);
  std::vector<std::vector<int32_t>> elements;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const int32_t v0 = j * (n + 1) + i, v1 = v0 + 1, v2 = v0 + n + 1, v3 = v2 + 1;
      if (i < n / 2)
        elements.push_back({v0, v1, v2, v3});
      else
        elements.push_back({v0, v1, v3}), elements.push_back({v0, v2, v3});
    }
  return Grid::mixed(coords, elements);
}

int main(int argc, char* argv[])
{
  try {
    const int n = argc > 1 ? std::atoi(argv[1]) : 128;
    const double diffusion = 1;
    const GenericFunction source(3, [](const double* x) {
      return M_PI_2 * M_PI * std::cos(M_PI_2 * x[0]) * std::cos(M_PI_2 * x[1]);
    });
    const GenericFunction exact_solution(
        3,
        /*evaluate=*/[](const double* x) { return std::cos(M_PI_2 * x[0]) * std::cos(M_PI_2 * x[1]); },
        /*name=*/"exact_solution",
        /*jacobian=*/
        [](const double* x, double* g) {
          const double pre = -0.5 * M_PI;
          g[0] = pre * std::sin(M_PI_2 * x[0]) * std::cos(M_PI_2 * x[1]);
          g[1] = pre * std::cos(M_PI_2 * x[0]) * std::sin(M_PI_2 * x[1]);
        });

    const bool mixed = argc > 3 && std::string(argv[3]) == "mixed";
    auto grid = mixed ? make_mixed_strip_grid(n)
                      : make_cube_grid(Geometry::quad, /*lower_left=*/-1., /*upper_right=*/1., /*num_elements=*/n);
    const auto& grid_view = grid.leaf_view();
    AllDirichletBoundaryInfo boundary_info;

    auto space = make_continuous_lagrange_space(grid_view, /*polorder=*/1);

    auto lhs_op = make_matrix_operator(space, Stencil::element);
    lhs_op.append(LocalElementIntegralBilinearForm(LocalLaplaceIntegrand(diffusion)));

    auto rhs_func = make_vector_functional(space);
    rhs_func.append(LocalElementIntegralFunctional(LocalProductIntegrand().with_ansatz(source)));

    auto dirichlet_constraints = make_dirichlet_constraints(space, boundary_info);

    // one device walk: matrix, functional and constraints
    lhs_op.append(rhs_func);
    lhs_op.append(dirichlet_constraints);
    lhs_op.assemble(/*use_tbb=*/true);

    auto solution = make_discrete_function(space);
    const auto stats = lhs_op.apply_inverse(lhs_op.rhs_device(), solution.dofs());

    auto h1_prod = make_bilinear_form(grid_view, solution - exact_solution);
    h1_prod += LocalElementIntegralBilinearForm(LocalLaplaceIntegrand(diffusion));
    auto l2_prod = make_bilinear_form(grid_view, solution - exact_solution);
    l2_prod += LocalElementIntegralBilinearForm(LocalProductIntegrand());
    const double h1 = std::sqrt(h1_prod.apply2()), l2 = std::sqrt(l2_prod.apply2());

    std::cout << "error in H^1 semi-norm: " << h1 << std::endl;
    std::cout << "error in L^2 norm:      " << l2 << std::endl;
    std::printf("{\"dofs\": %lld, \"iterations\": %lld, \"relres\": %.17g, \"h1_semi_error\": %.17g, "
                "\"l2_error\": %.17g, \"solve_ms\": %.6g}\n",
                (long long)space.size(), (long long)stats.iterations, stats.relres, h1, l2, stats.solve_ms);
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
