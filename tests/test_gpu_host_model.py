"""Host-model interface on the GPU: the reference's MONARCH mock
(test/monarch/mock_monarch.F90) with its simple configuration (the
consecutive A -> B -> C mechanism, interface_simple.json) on a 5 x 5 x 3
grid of 800 tracers, through include/camp_host_model.h.

* the gather is bit-exact against a numpy restatement of
  camp_monarch_interface.F90:455-471 (single-precision fields, the water
  product rounded to single precision before scaling);
* every step of every cell matches the closed-form solution of A -> B -> C
  applied to the gathered state (the mechanism is linear), after rounding
  to single precision;
* multi-cell equals single-cell integration within the reference's rule
  (almost_equal rel 1e-5, abs 1e-4; mock_monarch.F90:206-222), in per-cell
  and in block mode.
"""
import os

import numpy as np
import pytest

from camp_amd.core import MODE_BLOCK, MODE_PER_CELL
from camp_amd.host_model import MonarchInterface
from camp_amd.mechanisms import BOLTZMANN

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "monarch_simple")
NI, NJ, NK, NT = 5, 5, 3, 800
START, END = 100, 650
WV = 5
TIME_STEP = np.float32(1.6)
N_STEPS = 6  # i_time = 0 .. NUM_TIME_STEP
K_B = BOLTZMANN  # the constant the Arrhenius packer converts Ea with


def _mock_fields():
    """model_initialize (mock_monarch.F90:263-300), in single precision."""
    f32 = np.float32
    T = np.full((NI, NJ, NK), f32(300.614166259766), np.float32, order="F")
    P = np.full((NI, NK, NJ), f32(94165.7187500000), np.float32, order="F")
    for i in range(1, NI + 1):
        T[i - 1] = T[i - 1] + f32(0.1) * f32(i)
        P[i - 1] = P[i - 1] - f32(1 * i)
    for j in range(1, NJ + 1):
        T[:, j - 1] = T[:, j - 1] + f32(0.3) * f32(j)
        P[:, :, j - 1] = P[:, :, j - 1] - f32(3 * j)
    for k in range(1, NK + 1):
        T[:, :, k - 1] = T[:, :, k - 1] + f32(0.6) * f32(k)
        P[:, k - 1, :] = P[:, k - 1, :] - f32(6 * k)
    conc = np.zeros((NI, NJ, NK, NT), np.float32, order="F")
    water = np.zeros((NI, NJ, NK, WV), np.float32, order="F")
    water[..., WV - 1] = f32(0.01)
    rho = np.full((NI, NK, NJ), f32(1.225), np.float32, order="F")
    return T, P, conc, water, rho


def _interface(n_cells, **kw):
    return MonarchInterface(os.path.join(FIX, "config_simple.json"),
                            os.path.join(FIX, "interface_simple.json"), START, END, n_cells,
                            (NI, NJ, NK, NT), WV, **kw)


def _gather_np(hm, T, P, conc, water, rho):
    """camp_monarch_interface.F90:455-471 over the whole grid, cell z =
    (k-1)*NI*NJ + (j-1)*NI + (i-1)."""
    ns = hm.mech.n_state
    st = np.zeros((NI * NJ * NK, ns))
    env = np.zeros((NI * NJ * NK, 2))
    for k in range(NK):
        kf = NK - 1 - k
        for j in range(NJ):
            for i in range(NI):
                z = k * NI * NJ + j * NI + i
                for mid, cid in zip(hm.map_monarch_id, hm.map_camp_id):
                    st[z, cid - 1] = np.float64(conc[i, j, kf, mid - 1])
                wr = np.float32(water[i, j, kf, WV - 1]) * np.float32(rho[i, k, j])
                assert wr.dtype == np.float32
                st[z, hm.gas_phase_water_id - 1] = np.float64(wr) * 1.0e9
                env[z] = (np.float64(T[i, j, kf]), np.float64(P[i, k, j]))
    return st, env


def _exact(st, env, dt):
    """A -> B -> C (Arrhenius, A = 12 and 13, Ea = 1e-20 and 2e-20 J) over
    dt from the gathered state."""
    T = env[:, 0]
    k1 = 12.0 * np.exp(-1.0e-20 / K_B / T)
    k2 = 13.0 * np.exp(-2.0e-20 / K_B / T)
    a0, b0, c0 = st[:, 0], st[:, 1], st[:, 2]
    e1, e2 = np.exp(-k1 * dt), np.exp(-k2 * dt)
    a = a0 * e1
    b = b0 * e2 + a0 * k1 / (k2 - k1) * (e1 - e2)
    c = c0 + (a0 - a) + (b0 - b)
    return np.stack([a, b, c], axis=1)


def _run(n_cells, mode, check=False):
    T, P, conc, water, rho = _mock_fields()
    hm = _interface(n_cells, mode=mode)
    hm.get_init_conc(conc, water, WV, rho)
    assert hm.map_monarch_id.tolist() == [101, 100, 102]  # A, B, C (interface_simple.json)
    worst = 0.0
    for step in range(N_STEPS):
        if check:
            st, env = hm.gather(1, NI, 1, NJ, T, conc, water, WV, rho, P)
            st_np, env_np = _gather_np(hm, T, P, conc, water, rho)
            assert np.array_equal(st, st_np) and np.array_equal(env, env_np), step
        before = conc.copy()
        assert hm.integrate(TIME_STEP, 1, NI, 1, NJ, T, conc, water, WV, rho, P) == 0
        if check:
            want = _exact(st_np, env_np, float(TIME_STEP))
            got = np.zeros_like(want)
            for k in range(NK):
                kf = NK - 1 - k
                for j in range(NJ):
                    for i in range(NI):
                        z = k * NI * NJ + j * NI + i
                        got[z] = [conc[i, j, kf, m - 1] for m in hm.map_monarch_id]
            err = np.abs(got - want) / (1e-5 * np.abs(want) + 1e-9)
            worst = max(worst, float(err.max()))
            # tracers the map does not name are untouched
            other = np.ones(NT, bool)
            other[hm.map_monarch_id - 1] = False
            assert np.array_equal(conc[..., other], before[..., other])
    hm.close()
    return conc, worst


@pytest.mark.parametrize("mode", [MODE_PER_CELL, MODE_BLOCK])
def test_mock_monarch_multi_cell_equals_single_cell(mode):
    multi, worst = _run(NI * NJ * NK, mode, check=True)
    print("mode", mode, "worst closed-form error (units of 1e-5 rel, 1e-9 abs): %.3g" % worst)
    assert worst <= 1.0
    single, _ = _run(1, mode)
    a = multi[..., START - 1:END].astype(np.float64)
    b = single[..., START - 1:END].astype(np.float64)
    assert np.all(np.abs(a - b) <= 1e-5 * np.abs(b) + 1e-4)
    if mode == MODE_PER_CELL:
        print("per-cell mode: multi-cell and single-cell bitwise equal:", np.array_equal(a, b))


def test_unmapped_species_carry_over_in_multi_cell_mode(tmp_path):
    """A multi-cell interface keeps each cell's CAMP state on the GPU: with
    C left out of the map, C keeps the value set_state gave it through the
    gather, and after a step it has grown by what A and B lost, cell by
    cell; a single-cell interface clears the state before each cell."""
    import json
    cfg = json.load(open(os.path.join(FIX, "interface_simple.json")))
    del cfg["monarch-data"][0]["gas-phase species"]["C"]
    path = tmp_path / "interface_ab.json"
    path.write_text(json.dumps(cfg))
    T, P, conc, water, rho = _mock_fields()
    hm = MonarchInterface(os.path.join(FIX, "config_simple.json"), str(path), START, END,
                          NI * NJ * NK, (NI, NJ, NK, NT), WV)
    hm.get_init_conc(conc, water, WV, rho)
    n = NI * NJ * NK
    st0 = np.zeros((n, 3))
    st0[:, 2] = np.arange(n) + 7.0
    hm.set_state(st0)
    st, env = hm.gather(1, NI, 1, NJ, T, conc, water, WV, rho, P)
    assert np.array_equal(st[:, 2], st0[:, 2])
    assert hm.integrate(TIME_STEP, 1, NI, 1, NJ, T, conc, water, WV, rho, P) == 0
    st1, _ = hm.gather(1, NI, 1, NJ, T, conc, water, WV, rho, P)
    want = _exact(st, env, float(TIME_STEP))[:, 2]
    assert np.all(np.abs(st1[:, 2] - want) <= 1e-5 * np.abs(want))
    hm.close()
    one = MonarchInterface(os.path.join(FIX, "config_simple.json"), str(path), START, END, 1,
                           (NI, NJ, NK, NT), WV)
    one.set_state(np.full((1, 3), 5.0))
    st, _ = one.gather(1, NI, 1, NJ, T, conc, water, WV, rho, P)
    assert np.all(st[:, 2] == 0.0)
    one.close()
