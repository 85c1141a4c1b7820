"""What per-node smoother blocks can gain inside the multigrid preconditioner
(CPU, the FP64 model oracle/mg_model.py, which tests/test_gpu_mg.py pins to
the device V-cycle).

The learned network predicts per-node 3x3 SPD blocks.  Inside the V-cycle they
replace block Jacobi as the finest level's Chebyshev smoother.  Here the
blocks are not predicted but optimised directly through the V-cycle's error
propagation on the very problem they are measured on -- the best any network
could do there.  The V-cycle's convergence factor and PCG iteration count
barely move (measured 0.614 -> 0.601 and 23 -> 22 on this beam; 0.609 ->
0.595 and 21 -> 20 on the 64x16x16 beam with three levels), because what
limits this multigrid is its coarse-grid correction, not its smoother's
per-node blocks; so no per-node block network can make the learned V-cycle
meaningfully faster than block-Jacobi smoothing."""
import numpy as np
import pytest

from oracle import mg_model as MM
from solidfem_amd.mesh import BeamMesh, beam_hierarchy, lame

torch = pytest.importorskip("torch")


def hierarchy(grid):
    m = BeamMesh(*grid)
    fixed = m.fixed_mask()
    mu, lam = lame(1e6, 0.45)
    return m, MM.Hierarchy(m, fixed, beam_hierarchy(m), mu, lam)


def test_model_vcycle_symmetric_and_contracting():
    m, H = hierarchy((16, 4, 4))
    lms = [H.lmax(l, H.jacobi[l]) for l in range(H.nlev - 1)]
    rng = np.random.default_rng(2)
    n = H.K[0].shape[0]
    a, b = (torch.tensor(rng.standard_normal((n, 1))) for _ in range(2))
    Va, Vb = H.vcycle(a, H.jacobi, lms), H.vcycle(b, H.jacobi, lms)
    s1, s2 = float((b * Va).sum()), float((a * Vb).sum())
    assert abs(s1 - s2) <= 1e-10 * np.sqrt(float((a * Va).sum()) * float((b * Vb).sum()))
    assert float((a * Va).sum()) > 0
    # a Chebyshev-smoothed V-cycle with an exact coarsest solve contracts
    assert H.factor(H.jacobi, lms) < 0.7
    # with the exact inverse as the "V-cycle" PCG takes one step
    assert H.pcg_iterations(H.jacobi, lms, rng.standard_normal(n)) < 40


def test_optimised_smoother_blocks_gain_little():
    m, H = hierarchy((32, 8, 8))
    lms = [H.lmax(l, H.jacobi[l]) for l in range(H.nlev - 1)]
    rng = np.random.default_rng(0)
    rhs = rng.standard_normal(H.K[0].shape[0])
    f_bj = H.factor(H.jacobi, lms)
    it_bj = H.pcg_iterations(H.jacobi, lms, rhs)
    with torch.no_grad():
        loss_bj = float(MM.two_cycle_loss(H, H.jacobi, lms, MM.slow_errors(H, H.jacobi, lms)))
    M0, lmo = MM.optimise_smoother_blocks(H, steps=150)
    Ms = [M0] + H.jacobi[1:]
    with torch.no_grad():
        loss_opt = float(MM.two_cycle_loss(H, Ms, lmo, MM.slow_errors(H, Ms, lmo)))
    f_opt = H.factor(Ms, lmo)
    it_opt = H.pcg_iterations(Ms, lmo, rhs)
    print(f"slow-mode loss {loss_bj:.4f} -> {loss_opt:.4f}; V-cycle factor {f_bj:.4f} -> {f_opt:.4f}; "
          f"PCG iterations {it_bj} -> {it_opt}")
    # the optimisation does lower its own loss (the slow modes' two-cycle reduction) ...
    assert loss_opt < 0.97 * loss_bj
    # ... yet the V-cycle converges at most a few percent faster
    assert f_opt > 0.95 * f_bj
    assert it_opt >= it_bj - 1
