"""svimx_velocity_static, the entry point that allocates nothing, against
svimx_velocity (workspace form) through the C ABI: bit-identical velocities and
statistics, on two calls in a row with the second on a smaller band, for an SVI
and a curvature law, and with the same results when the calls repeat."""

import ctypes as C

import numpy as np
import pytest
import torch

from tests.conftest import sphere_phi

pytestmark = pytest.mark.gpu


def _job(R, h, law):
    """A sphere level set with the velocity inputs prepared (N_r or curvature pairs)."""
    from svierosion_amd import _abi
    from svierosion_amd.evolve import ErosionSimulation
    from svierosion_amd.grid import Grid
    from svierosion_amd.levelset import LevelSet
    phi, origin, h, shape = sphere_phi(R, h, 0.3 + 12 * h)
    ls = LevelSet(torch.from_numpy(phi.astype(np.float32)).cuda(), Grid(origin, h, shape), width=6)
    sim = ErosionSimulation(ls, law)
    sim._prepare(ls.cur, True)       # fractions, N_r / curvatures, velocity (workspace form)
    torch.cuda.synchronize()
    return ls, sim, _abi


def _both(ls, sim, _abi, results):
    """Velocity and statistics by both entry points on the same inputs."""
    st = _abi.stream()
    V_ws = torch.empty_like(sim.V)
    V_st = torch.full_like(sim.V, float("nan"))
    args = (st, _abi.ptr(ls.flat), _abi.ptr(ls._band), ls.nb, *ls.dims, ls.h, sim._law_abi, _abi.ptr(sim.field),
            _abi.ptr(sim.k12), sim.cutoff[0], sim.cutoff[1], 1)
    _abi.check(_abi.lib.svimx_velocity(*args, _abi.ptr(V_ws), _abi.ptr(ls.ws), ls.ws.numel()))
    _abi.check(_abi.lib.svimx_velocity_static(*args, _abi.ptr(V_st), _abi.ptr(results)))
    torch.cuda.synchronize()
    lo, hi = _abi.R_VMAX, _abi.R_DMAX_ALL + 1
    return V_ws, V_st, _abi.results_view(ls.ws)[lo:hi].clone(), results[lo:hi].clone()


@pytest.mark.parametrize("kind", ["svi", "mean"])
def test_static_velocity_matches_workspace_form(device, kind):
    from svierosion_amd.flows import MeanCurvatureLaw, SVIPowerLaw
    law = SVIPowerLaw(0.3, 1, 1, 0.93) if kind == "svi" else MeanCurvatureLaw(1.0)
    results = torch.full((32,), float("nan"), dtype=torch.float64, device="cuda")
    big = _job(1.0, 0.015, law)         # ~670k band nodes: the full 2048-block reduction grid
    small = _job(0.5, 0.05, law)        # a much smaller band (fewer blocks, stale partials beyond)
    assert big[0].nb > 256 * 2048 and small[0].nb < big[0].nb // 10
    for ls, sim, abi in (big, small, big):
        V_ws, V_st, r_ws, r_st = _both(ls, sim, abi, results)
        nb = ls.nb
        print(f"{kind}: {nb} band nodes, statistics {r_st.cpu().numpy()[:4]}", flush=True)
        assert torch.equal(V_ws[:nb], V_st[:nb])
        assert torch.equal(r_ws, r_st), (r_ws, r_st)


def test_static_velocity_rejects_bad_arguments(device):
    from svierosion_amd import _abi
    from svierosion_amd.flows import SVIPowerLaw
    law = SVIPowerLaw(0.3, 1, 1, 0.93).to_abi()
    st = _abi.stream()
    # statistics without a results array, and a band beyond the int32 list
    rc = _abi.lib.svimx_velocity_static(st, None, None, 0, 8, 8, 8, 0.1, C.byref(law), None, None, 4.0, 5.5, 1,
                                        None, None)
    assert rc == _abi.EBADARG and "results" in _abi.lib.svimx_last_error().decode()
    res = torch.zeros(32, dtype=torch.float64, device="cuda")
    rc = _abi.lib.svimx_velocity_static(st, None, None, 1 << 31, 8, 8, 8, 0.1, C.byref(law), None, None, 4.0, 5.5,
                                        1, None, _abi.ptr(res))
    assert rc == _abi.EBADARG
