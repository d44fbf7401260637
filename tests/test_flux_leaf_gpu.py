"""fcfd_convective_flux (the leaf flux functions over a batch of faces)
against the oracle's convective_face_flux of the same two states
(orc_flux_probe): the flux and weightL bit for bit, for the five schemes,
over random face frames and states from Mach 0 to 2.5 in every direction,
including equal states and the supersonic upwind branches; Steger-Warming
with and without its wall-distance dissipation arrays."""
import numpy as np
import pytest

from freecfd_amd import _abi as A
from freecfd_amd import cases

pytestmark = pytest.mark.gpu

FLUXES = [A.FLUX_ROE, A.FLUX_VAN_LEER, A.FLUX_AUSM_PLUS_UP, A.FLUX_SD_SLAU, A.FLUX_SW]


def _faces(n, seed, R, gamma=1.4):
    rng = np.random.default_rng(seed)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    t1 = np.cross(nrm, rng.normal(size=(n, 3)))
    t1 /= np.linalg.norm(t1, axis=1)[:, None]
    W = []
    for _ in range(2):
        T = 250.0 + 100.0 * rng.random(n)
        a = np.sqrt(gamma * R * T)
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        M = 2.5 * rng.random(n)
        V = (M * a)[:, None] * d
        p = 1e5 * (0.5 + rng.random(n))
        W.append(np.column_stack([p, V, T]))
    WL, WR = W
    WR[:8] = WL[:8]                      # equal states: the consistency branch
    WL[8:16, 1:4] = 3.0 * np.sqrt(gamma * R * WL[8:16, 4])[:, None] * nrm[8:16]    # supersonic along +n
    WR[8:16, 1:4] = 2.5 * np.sqrt(gamma * R * WR[8:16, 4])[:, None] * nrm[8:16]
    WL[16:24, 1:4] = -3.0 * np.sqrt(gamma * R * WL[16:24, 4])[:, None] * nrm[16:24]   # supersonic along -n
    WR[16:24, 1:4] = -2.5 * np.sqrt(gamma * R * WR[16:24, 4])[:, None] * nrm[16:24]
    area = 0.1 + rng.random(n)
    return WL, WR, nrm, t1, area


@pytest.mark.parametrize("flux", FLUXES)
def test_convective_flux_matches_the_oracle_bitwise(flux):
    from oracle.oracle import Oracle
    from freecfd_amd.ns import NavierStokes
    cs = cases.wedge_case(8, 4)
    print("flux", flux, flush=True)
    o = Oracle(cs.raw, cs.opt)
    ns = NavierStokes(cs.raw, cs.opt)
    n = 160
    WL, WR, nrm, t1, area = _faces(n, 100 + flux, cases.gas_R(28.97))
    Fg, wg = ns.convective_flux(flux, WL.T, WR.T, nrm.T, t1.T, area)
    Fg, wg = Fg.cpu().numpy(), wg.cpu().numpy()
    for i in range(n):
        Fo, wo = o.flux_probe(flux, WL[i], WR[i], nrm[i], t1[i], area=area[i], weight=True)
        assert np.array_equal(Fg[:, i], Fo), (flux, i, Fg[:, i], Fo)
        assert wg[i] == wo, (flux, i, wg[i], wo)
    if flux == A.FLUX_SW:
        # explicit dissipation arrays equal to the defaults give the same bits;
        # a near-wall face with other values differs
        Fd, _ = ns.convective_flux(flux, WL.T, WR.T, nrm.T, t1.T, area, diss=np.ones(n), wall_dist=np.full(n, 1e20))
        assert np.array_equal(Fd.cpu().numpy(), Fg)
        Fw, _ = ns.convective_flux(flux, WL.T, WR.T, nrm.T, t1.T, area, diss=np.full(n, 0.3), wall_dist=np.zeros(n))
        assert not np.array_equal(Fw.cpu().numpy(), Fg)
    ns.close()


def test_convective_flux_argument_errors():
    from freecfd_amd.ns import NavierStokes
    cs = cases.wedge_case(8, 4)
    ns = NavierStokes(cs.raw, cs.opt)
    WL, WR, nrm, t1, area = _faces(4, 1, cases.gas_R(28.97))
    with pytest.raises(RuntimeError):
        ns.convective_flux(9, WL.T, WR.T, nrm.T, t1.T, area)
    with pytest.raises(ValueError):
        ns.convective_flux(A.FLUX_ROE, WL, WR.T, nrm.T, t1.T, area)
    F, w = ns.convective_flux(A.FLUX_ROE, np.zeros((5, 0)), np.zeros((5, 0)), np.zeros((3, 0)), np.zeros((3, 0)), np.zeros(0))
    assert F.shape == (5, 0) and w.shape == (0,)
    ns.close()
