"""Batched solves (union of independent grids) and the adjoint gradient on the HIP path.

Batch: every instance of a union grid is assembled in the same launches and solved by the same PCG
launches with its own scalars and stopping test.  Since an instance's rows are partitioned into
blocks exactly as in a plain operator with the same launch shapes, its iterates are the same
floating-point operations in the same order: the batched solution equals the plain per-instance
GPU solution bit for bit, and matches the oracle to 1e-10.

Adjoint: dJ/d(log kappa_e) = -kappa_e lambda_e^T Khat_e u_e through torch.autograd, checked against
the oracle's gradient (1e-8 relative) and central finite differences of J (h = 1e-6).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from problems import hgdt_problem

pytestmark = pytest.mark.gpu


def _batch_problem(hgdt, grids, kappa, f):
    space = hgdt.ContinuousLagrangeSpace(grids, 1)
    op = hgdt.make_matrix_operator(space, hgdt.Stencil.element)
    op.append(hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa)))
    rhs = hgdt.make_vector_functional(space)
    rhs.append(hgdt.LocalElementIntegralFunctional(hgdt.LocalProductIntegrand().with_ansatz(f)))
    op.append(rhs)
    op.append(hgdt.make_dirichlet_constraints(space))
    return space, op


def test_batched_equals_plain_and_oracle(hgdt_mod):
    print("case batch of uneven Q1 grids n = 5, 9, 17, 31, 64", flush=True)
    F = hgdt_mod.fields
    sizes = [5, 9, 17, 31, 64]
    parts = [hgdt_mod.Grid.structured("quad", [0, 0], [1, 1], [n, n]) for n in sizes]
    union = hgdt_mod.Grid.union(parts)
    assert union.num_instances == len(sizes)
    ne = union.num_elements
    kappa = F.lognormal(21, ne)
    f = F.uniform(22, ne)
    space, op = _batch_problem(hgdt_mod, union, torch.from_numpy(kappa).cuda(), torch.from_numpy(f).cuda())
    op.assemble()
    xb = op.apply_inverse("rhs", use_graph=True, check_every=8).cpu().numpy()
    its, rels = op.instance_stats()
    eoff = union.instance_elements()
    doff = space.instance_dofs()
    assert np.all(rels <= 1e-10)
    # the small instances finish long before n = 64: later chunks run over the unfinished blocks only
    assert op.last_solve["compactions"] >= 1
    for i, n in enumerate(sizes):
        kap_i, f_i = kappa[eoff[i]:eoff[i + 1]], f[eoff[i]:eoff[i + 1]]
        # plain GPU operator of the same instance, same launch shapes
        g = hgdt_mod.Grid.structured("quad", [0, 0], [1, 1], [n, n])
        s1, op1 = _batch_problem(hgdt_mod, g, torch.from_numpy(kap_i).cuda(), torch.from_numpy(f_i).cuda())
        for k in (3, 4):
            op1.set_launch(k, *op.get_launch(k))
        op1.assemble()
        x1 = op1.apply_inverse("rhs").cpu().numpy()
        assert op1.last_solve["iterations"] == its[i]
        assert np.array_equal(x1, xb[doff[i]:doff[i + 1]]), f"instance {i} differs from its plain solve"
        # oracle
        m, s, oop = None, None, None
        mo = O.Mesh.structured(O.QUAD, [n, n], [0, 0], [1, 1])
        so = O.Space(mo, True, 1)
        oop = O.Operator(so, O.STENCIL_ELEMENT)
        oop.append_laplace(kap_i)
        oop.append_source(f_i)
        oop.set_dirichlet(True)
        oop.assemble(1)
        rp, cols = oop.pattern()
        xo, ito, relo, rc = O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=1e-10)
        rel = np.linalg.norm(xb[doff[i]:doff[i + 1]] - xo) / np.linalg.norm(xo)
        print(f"  instance {i} n={n}: iters {its[i]} (oracle {ito}) rel L2 {rel:.2e}")
        assert rel <= 1e-10 and abs(its[i] - ito) <= 2


def test_union_of_union_rejected(hgdt_mod):
    g = hgdt_mod.Grid.structured("quad", [0, 0], [1, 1], [4, 4])
    u = hgdt_mod.Grid.union([g, g])
    with pytest.raises(hgdt_mod.wrong_input_given):
        hgdt_mod.Grid.union([u, g])


def _oracle_gradient(n, logk, u_obs_logk):
    m = O.Mesh.structured(O.TET, [n, n, n], [0, 0, 0], [1, 1, 1])
    s = O.Space(m, True, 1)

    def solve(lk):
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_laplace(np.exp(lk))
        op.append_source(1.0)
        op.set_dirichlet(True)
        op.assemble(1)
        rp, cols = op.pattern()
        x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-12, max_iter=100000)
        return x, (rp, cols, op.values())

    u_obs, _ = solve(u_obs_logk)
    u, (rp, cols, vals) = solve(logk)
    lam, _, _, _ = O.pcg_jacobi(rp, cols, vals, u - u_obs, rtol=1e-12, max_iter=100000)
    return O.adjoint_log_kappa_gradient(s, np.exp(logk), u, lam), u_obs, solve


def test_adjoint_gradient_matches_oracle_and_fd(hgdt_mod):
    print("case adjoint P1 tets 5^3 x 6, two instances", flush=True)
    from hgdt.autograd import DiffusionProblem, solve_log_diffusion

    F = hgdt_mod.fields
    n, ninst = 5, 2
    part = hgdt_mod.Grid.structured("tet", [0, 0, 0], [1, 1, 1], [n, n, n])
    union = hgdt_mod.Grid.union([part] * ninst)
    ne1 = part.num_elements
    logk = F.normal(31, ne1 * ninst)
    logk_obs = F.normal(32, ne1 * ninst)
    prob = DiffusionProblem(union, source=1.0, rtol=1e-12)
    doff = prob.space.instance_dofs()
    u_obs = torch.cat([torch.from_numpy(_oracle_gradient(n, logk[i * ne1:(i + 1) * ne1],
                                                          logk_obs[i * ne1:(i + 1) * ne1])[1])
                       for i in range(ninst)]).cuda()
    theta = torch.tensor(logk, device="cuda", requires_grad=True)
    u = solve_log_diffusion(theta, prob)
    J = 0.5 * torch.sum((u - u_obs) ** 2)
    J.backward()
    g = theta.grad.detach().cpu().numpy()
    for i in range(ninst):
        go, _, solve = _oracle_gradient(n, logk[i * ne1:(i + 1) * ne1], logk_obs[i * ne1:(i + 1) * ne1])
        gi = g[i * ne1:(i + 1) * ne1]
        rel = np.abs(gi - go).max() / np.abs(go).max()
        print(f"  instance {i}: gradient vs oracle max rel {rel:.2e}")
        assert rel <= 1e-8
    # central differences on the GPU path itself (J through the batched forward solve)
    rng = np.random.default_rng(7)
    h = 1e-6
    gmax = np.abs(g).max()
    # J(theta +- h e) differ by ~h |g|: solve tightly so that solver error stays below the difference
    prob.rtol = 1e-13
    for e in rng.choice(ne1 * ninst, 6, replace=False):
        vals = []
        for sgn in (1, -1):
            lk = logk.copy()
            lk[e] += sgn * h
            with torch.no_grad():
                up = solve_log_diffusion(torch.tensor(lk, device="cuda"), prob)
                vals.append(float(0.5 * torch.sum((up - u_obs) ** 2)))
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"  element {e}: adjoint {g[e]:.6e} fd {fd:.6e}")
        assert abs(fd - g[e]) <= 1e-5 * gmax


def test_adjoint_backward_after_a_second_forward(hgdt_mod):
    """two forward solves on one DiffusionProblem (kappa_1, then kappa_2) and the backward of the first:
    its gradient must be the oracle's at kappa_1, not at the kappa the operator was last assembled with;
    the second graph's gradient must then be the oracle's at kappa_2"""
    print("case adjoint: forward(k1), forward(k2), backward(first), backward(second)", flush=True)
    from hgdt.autograd import DiffusionProblem, solve_log_diffusion

    F = hgdt_mod.fields
    n = 4
    grid = hgdt_mod.Grid.structured("tet", [0, 0, 0], [1, 1, 1], [n, n, n])
    ne = grid.num_elements
    lk1, lk2, lk_obs = F.normal(61, ne), F.normal(62, ne), F.normal(63, ne)
    g1, u_obs, _ = _oracle_gradient(n, lk1, lk_obs)
    g2, _, _ = _oracle_gradient(n, lk2, lk_obs)
    prob = DiffusionProblem(grid, source=1.0, rtol=1e-12)
    u_obs = torch.from_numpy(u_obs).cuda()
    t1 = torch.tensor(lk1, device="cuda", requires_grad=True)
    t2 = torch.tensor(lk2, device="cuda", requires_grad=True)
    J1 = 0.5 * torch.sum((solve_log_diffusion(t1, prob) - u_obs) ** 2)
    J2 = 0.5 * torch.sum((solve_log_diffusion(t2, prob) - u_obs) ** 2)
    J1.backward()
    assert prob.reassemblies == 1
    J2.backward()
    assert prob.reassemblies == 2
    for t, go, label in ((t1, g1, "first"), (t2, g2, "second")):
        g = t.grad.detach().cpu().numpy()
        rel = np.abs(g - go).max() / np.abs(go).max()
        print(f"  {label} graph: gradient vs oracle max rel {rel:.2e}")
        assert rel <= 1e-8, label
    # the two gradients differ, so the check above can tell them apart
    assert np.abs(g1 - g2).max() > 1e-3 * np.abs(g1).max()
