"""CPU oracle: one time step (the Newton loop of main_helper::advance).

Pinned by physics that does not depend on the implementation:
  * free fall of a rest-shaped sheet (contact off): every vertex moves by
    dt^2 g up to the air-drag factor m / (m + rho_air A) along the normal;
  * pinned vertices stay exactly at their fix positions (Dirichlet);
  * a sheet dropped on a sphere never penetrates it and never intersects
    itself (the post-step check), and the accumulated toi bookkeeping
    shortens dt only on contact.
"""
import numpy as np

import oracle
from ppf_amd import scenes


def test_free_fall_closed_form():
    sc = scenes.drape_scene(res=8, height=0.5)
    st = oracle.OracleStepper(sc)
    x0 = sc["mesh"]["x"].copy()
    p = scenes.step_params(disable_contact=1)
    r = st.advance(p)
    assert r["status"] == 0 and r["toi_advanced"] == 1.0
    dt = p["dt"]
    d = sc["mesh"]["x"] - x0
    m, a = sc["shell"]["vmass"], sc["shell"]["varea"]
    # air drag on the (vertical) normal component: m y'' = -m g - rho a y'' ...
    expect = -9.8 * dt * dt * m / (m + 1e-3 * a)
    assert np.allclose(d[:, 1], expect, rtol=2e-3), (d[:, 1][:4], expect[:4])
    # no tangential motion beyond fp32 rounding of coordinates of size ~0.5
    assert np.abs(d[:, [0, 2]]).max() < 2.0 ** -22


def test_pins_stay_and_no_penetration():
    res = 10
    sc = scenes.drape_scene(res=res, height=0.004, pins=[0, res - 1])
    fix = sc["fix_pos"].copy()
    st = oracle.OracleStepper(sc)
    p = scenes.step_params()
    shortened = 0
    for k in range(20):
        r = st.advance(p)
        assert r["status"] == 0, (k, r)
        p["prev_dt"] = r["dt"]
        shortened += r["dt"] < p["dt"]
        x = sc["mesh"]["x"]
        assert np.array_equal(x[[0, res - 1]], fix[[0, res - 1]])
        # never inside the sphere (radius 0.3; contact gap 1e-3, offset 0)
        dist = np.linalg.norm(x, axis=1)
        assert dist.min() > 0.3 - 1e-4, (k, dist.min())
    assert shortened > 0        # contact shortened some steps


def test_kinematic_pins_rewind_to_the_integrated_fraction():
    # Pins slide along x; each step they are aimed at the step's end. When the
    # line search shortens the step to toi_advanced < 1, the pins are walked
    # back to position - (1 - toi_advanced) * step_delta before the final
    # iteration (main.cu:513-545), so they end there exactly.
    res = 10
    pins = [0, res - 1]
    sc = scenes.drape_scene(res=res, height=0.004, pins=pins)
    sc["fix_delta"] = np.zeros_like(sc["fix_pos"])
    st = oracle.OracleStepper(sc)
    p = scenes.step_params()
    v = np.float32(0.05 * p["dt"])
    rewound = 0
    for k in range(16):
        start = sc["mesh"]["x"][pins].copy()
        aim = start + np.array([v, 0, 0], np.float32)
        sc["fix_pos"][pins] = aim
        sc["fix_delta"][pins] = [v, 0, 0]
        st.delta[:] = sc["fix_delta"].reshape(-1)
        r = st.advance(p)
        assert r["status"] == 0, (k, r)
        p["prev_dt"] = r["dt"]
        x = sc["mesh"]["x"][pins]
        ta = np.float32(r["toi_advanced"])
        if ta < 1.0:
            back = np.float32(1.0) - ta
            expect = aim - back * np.array([v, 0, 0], np.float32)
            rewound += 1
        else:
            expect = aim
        assert np.array_equal(x, expect.astype(np.float32)), (k, x, expect)
    assert rewound > 0
