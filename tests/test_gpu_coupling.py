"""The same pins of the coupled right-hand side as test_oracle_coupling.py,
run on the GPU library through the C-ABI: the linear growth rate of the
leading eigenmode about conduction, the kinetic-energy budget and the
mean-heat balance, each computed with numpy from the fields libfdns returns.
"""
from __future__ import annotations

import numpy as np
import pytest

from test_oracle_coupling import (BUD, LIN, budgets, eigen_state, energy_err, growth_rate, heat_err, linear_mode,
                                  smooth_state)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lin_ns():
    return linear_mode(LIN["N"], LIN["m"], LIN["Ra"], LIN["Pr"], "NS")


def gpu_factory(fd, N, Nm, Ra, Pr, scheme, solver):
    return lambda dt: fd.Run(N, Nm, Ra, Pr, dt, scheme, restart=True, solver=solver)


@pytest.mark.parametrize("solver", ["lu", "inverse"])
@pytest.mark.parametrize("scheme,dt,steps", [("CNAB2", 1e-4, 40), ("SBDF3", 1e-4, 40), ("ARS443", 1e-4, 40),
                                             ("BPR353", 1e-4, 40), ("BHR553", 1e-4, 40), ("RK4", 2e-6, 400)])
def test_gpu_linear_growth_rate(fd, orc, lin_ns, scheme, dt, steps, solver):
    if scheme == "RK4" and solver == "inverse":
        pytest.skip("the explicit RK path has no implicit solves")
    N, Nm, m, Ra, Pr = LIN["N"], LIN["Nm"], LIN["m"], LIN["Ra"], LIN["Pr"]
    lam = lin_ns[0]
    state, th = eigen_state(orc, N, Nm, m, lin_ns)
    g = growth_rate(gpu_factory(fd, N, Nm, Ra, Pr, scheme, solver), state, th, m, dt, steps)
    assert abs(g - lam) < 1e-5 * abs(lam), (scheme, solver, g, lam)


@pytest.mark.parametrize("scheme,dt,warm", [("CNAB2", 2e-5, 200), ("SBDF3", 2e-5, 200), ("ARS443", 2e-5, 200),
                                            ("BHR553", 2e-5, 200), ("RK4", 2e-6, 2000)])
def test_gpu_energy_and_heat_budgets(fd, orc, scheme, dt, warm):
    N, Nm, Ra, Pr = BUD["N"], BUD["Nm"], BUD["Ra"], BUD["Pr"]
    run = fd.Run(N, Nm, Ra, Pr, dt, scheme, restart=True)
    run.set_state(*smooth_state(orc, N, Nm))
    run.step(warm)
    dke, buo, dis, dh, flux = budgets(run, N, Ra, Pr, dt)
    run.close()
    assert buo.min() > 0.2 * dis.max()
    assert energy_err(dke, buo, dis) < 1e-3, (scheme, dke, buo - dis)
    assert heat_err(dh, flux) < 1e-3, (scheme, dh, flux)
