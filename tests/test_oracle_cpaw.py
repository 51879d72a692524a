"""The oracle's unsplit Riemann driver (oracle/unsplit.c) against the
reference's own known answers: the circularly polarised Alfven wave test
(Makefile:351-366 runs it at 128 x 64 and expects L2 = 0.00102726;
problems/cpaw/problem.par:2-13 lists the norm at every resolution).  The
L2 error of B after one period depends on the whole scheme -- PLM with the
van Leer limiter, HLLD with the GLM psi flux, the unsplit RK2 stages, the
two half-step GLM dampings, c_h = cfl_glm*min(dl)/dt and the CFL timestep
with its clamp to tend -- and each printed digit is matched: a different
limiter, no psi damping, another cfl_glm or cfl all change the 8th digit
(the norm does not resolve rounding, so it pins the algorithm, not the
bits; the GPU path is held to the oracle's bits separately).
"""
import pytest

from piernik_amd import setups

SIZES = [(32, 16), (48, 24), (64, 32), (96, 48), (128, 64)]


@pytest.mark.parametrize("nx,ny", SIZES)
def test_cpaw_l2_matches_reference(oracle, nx, ny):
    cg, cfg, ini = setups.cpaw(nx, ny)
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    t, ns = oracle.run_unsplit(p, ob, setups.CPAW_TEND)
    assert t >= setups.CPAW_TEND and ns == 7 * nx // 4
    l2 = setups.cpaw_l2(cg.interior(ob.b), ini)
    assert f"{l2:.8f}" == f"{setups.CPAW_L2[(nx, ny)]:.8f}", l2


def test_cpaw_l2_catches_a_wrong_limiter(oracle):
    """The known answer is not met by a plausible mistake."""
    cg, cfg, ini = setups.cpaw(32, 16)
    cfg.limiter = cfg.limiter_b = "minmod"
    p = oracle.params_from_config(cfg)
    ob = oracle.OBlock.from_cg(cg)
    oracle.run_unsplit(p, ob, setups.CPAW_TEND)
    assert f"{setups.cpaw_l2(cg.interior(ob.b), ini):.8f}" != f"{setups.CPAW_L2[(32, 16)]:.8f}"
