"""A Mara driver script run twice in the reference's own Lua interpreter (oracle/_ref): once
on the reference's Mara (its CPU path) and once on the device library through the Lua
binding INTEGRATION.md shows (MaraGPU), with the same calls in the same order
(modules/MaraSim.lua:39-52, the dt = 0 priming step of modules/MaraSim.lua:91-94, then CFL
steps from get_timestep). The device must give the reference's states to the plain form's
full-step tolerances (L1 1e-10, 1e-7 at any zone), the same failed-zone counts and the
same time steps to 1e-12.
"""
import numpy as np
import pytest

from oracle import lua_ref

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (lua_ref.have_lua() and lua_ref.have_stub()),
                                 reason="oracle/_ref Lua front end not built")]

DRIVER = """
local M = {M}
M.start()
M.set_fluid('rmhd')
M.set_eos('gamma-law', 4/3)
M.set_advance('{advance}')
M.set_godunov('plm-split')
M.set_boundary('periodic')
M.set_riemann('{riemann}')
M.config_solver({{extrap = '{extrap}', theta = 2.0, IS = 'js96'}}, false)
M.set_domain({{0, 0, 0}}, {{1, 1, 1}}, {{{N}, {N}, {N}}}, 8, 3)
dts, errs = {{}}, {{}}
local dt = 0.0
for n = 1, {steps} do
  local kzps, nerr = M.advance({buf}, dt)
  errs[n] = nerr
  dt = M.get_timestep(0.4)
  dts[n] = dt
end
M.close()
"""


@pytest.mark.parametrize("extrap,riemann,advance", [("plm", "hlld", "rk2"), ("weno5", "hlld", "rk3")])
def test_driver_script_on_mara_and_on_the_device(gpu, extrap, riemann, advance):
    from test_gpu_parity import compare_states
    N, steps = 16, 4
    print(f"[lua] {extrap}/{riemann}/{advance} {N}^3, {steps} advances", flush=True)
    P0 = gpu.init_turbulence((N, N, N), v_rms=0.4, v_kmax=4.0, four_velocity=True, dB_rms=0.3,
                             A_kmax=3.0, seed=21)
    s = lua_ref.State(with_device=True)
    out = {}
    for M, buf in (("Mara", "Pref"), ("MaraGPU", "Pdev")):
        s.set_buffer(buf, P0)
        s.run(DRIVER.format(M=M, buf=buf, N=N, steps=steps, extrap=extrap, riemann=riemann,
                            advance=advance))
        out[M] = (s.get_buffer(buf, P0.shape), s.numbers("dts", steps), s.numbers("errs", steps))
    s.close()
    (Pr, dtr, er), (Pd, dtd, ed) = out["Mara"], out["MaraGPU"]
    assert np.array_equal(er, ed) and not er.any()
    assert np.all(np.abs(dtd - dtr) <= 1e-12 * dtr)
    assert not np.array_equal(Pr, P0)
    compare_states(Pd, Pr)
