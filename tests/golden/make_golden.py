"""Writes tests/golden/hertz_collision.json: restitution coefficient and contact
duration of a head-on sphere-sphere collision under the damped Hertz law,
obtained by integrating the contact ODE with scipy at tight tolerances.

The ODE is the normal-direction equation of motion of the relative overlap
delta(t) for two spheres with effective mass m*, effective radius R* and
effective modulus Y*:

    m* delta'' = -k delta^(3/2) - c delta^(1/4) delta'
    k = 4/3 Y* sqrt(R*)                       (Hertz, static force law)
    c = 2 sqrt(5/6) (-beta) sqrt(2 Y* sqrt(R*) m*),
    beta = ln e / sqrt(ln^2 e + pi^2)         (viscoelastic damping coefficient)

as stated by the reference's NormalModel<HERTZ> (normal_model_hertz.h:230-255)
and MODEL_PARAMS::createBetaEff (global_properties.cpp:525-544). This is the
physics the contact law must reproduce; the tests drive the oracle and the GPU
library through whole collisions (neighbor list, pair loop, Verlet integrator)
and compare the outgoing/incoming speed ratio and the contact time with it.

For e = 1 (beta = 0) the contact time also has Hertz's closed form
t_c = 2.9432... * delta_max / v0 with delta_max = (5 m* v0^2 / (4 k))^(2/5);
the generator checks its ODE solve against that before writing.

Run: python tests/golden/make_golden.py
"""
import json
import math
import os

import numpy as np
from scipy.integrate import quad, solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))


def collide(e, v0, r1, r2, rho, Y, nu):
    m1 = rho * 4.0 / 3.0 * math.pi * r1 ** 3
    m2 = rho * 4.0 / 3.0 * math.pi * r2 ** 3
    ms = m1 * m2 / (m1 + m2)
    Rs = r1 * r2 / (r1 + r2)
    Ys = 1.0 / ((1 - nu ** 2) / Y + (1 - nu ** 2) / Y)
    le = math.log(e)
    beta = le / math.sqrt(le * le + math.pi ** 2)
    k = 4.0 / 3.0 * Ys * math.sqrt(Rs)
    c = 2.0 * math.sqrt(5.0 / 6.0) * (-beta) * math.sqrt(2.0 * Ys * math.sqrt(Rs) * ms)

    def rhs(t, y):
        d, dd = y
        dp = max(d, 0.0)
        return [dd, (-k * dp ** 1.5 - c * dp ** 0.25 * dd) / ms]

    def leave(t, y):
        return y[0]
    leave.terminal = True
    leave.direction = -1
    tmax = 50.0 * (ms ** 2 / (Rs * Ys ** 2 * v0)) ** 0.2
    sol = solve_ivp(rhs, (0.0, tmax), [0.0, v0], events=leave, rtol=1e-12, atol=1e-20,
                    method="DOP853", max_step=tmax / 20000.0, first_step=tmax * 1e-9)
    # skip the t=0 root: the event fires only on a downward crossing after the peak
    tc = float(sol.t_events[0][0])
    vout = float(sol.y_events[0][0][1])
    dmax = float(np.max(sol.y[0]))
    return {"e": e, "v0": v0, "r1": r1, "r2": r2, "rho": rho, "Y": Y, "nu": nu,
            "cor": -vout / v0, "t_contact": tc, "delta_max": dmax, "m_eff": ms, "R_eff": Rs,
            "Y_eff": Ys}


def hertz_closed_form_tc(v0, ms, Rs, Ys):
    k = 4.0 / 3.0 * Ys * math.sqrt(Rs)
    dmax = (5.0 * ms * v0 ** 2 / (4.0 * k)) ** 0.4
    integral = quad(lambda x: 1.0 / math.sqrt(1.0 - x ** 2.5), 0.0, 1.0, limit=200)[0]
    return 2.0 * integral * dmax / v0, dmax


def main():
    cases = []
    for e in (1.0, 0.9, 0.5, 0.3, 0.1):
        for (v0, r1, r2) in ((1.0, 0.002, 0.002), (0.3, 0.0015, 0.001)):
            cases.append(collide(e, v0, r1, r2, 2500.0, 5e6, 0.45))
    for c in cases:
        if c["e"] == 1.0:
            tc, dmax = hertz_closed_form_tc(c["v0"], c["m_eff"], c["R_eff"], c["Y_eff"])
            assert abs(tc - c["t_contact"]) < 1e-6 * tc, (tc, c["t_contact"])
            assert abs(dmax - c["delta_max"]) < 1e-6 * dmax
            assert abs(c["cor"] - 1.0) < 1e-8
            c["t_contact_closed_form"] = tc
    with open(os.path.join(HERE, "hertz_collision.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f, indent=1)
    for c in cases:
        print(f"e={c['e']:.2f} v0={c['v0']} cor={c['cor']:.6f} tc={c['t_contact']:.6e}")


if __name__ == "__main__":
    main()
