"""Writes the golden fixtures of the oracle tests (run from the repository root:
python tests/golden/make_golden.py).  Every value comes from outside the code
under test:

* toro_table.json -- the five standard shock-tube tests and their exact
  star-region solutions as published in E. F. Toro, "Riemann Solvers and
  Numerical Methods for Fluid Dynamics" (3rd ed.), Tables 4.1 and 4.3
  (gamma = 1.4; values quoted to the book's 5-6 significant digits).
* kat_reference_test_kernel.json -- the reference's only test-like artefact,
  the unlaunched `test` kernel (kernel.cu:3095-3107): a uniform state
  rho = 1, p = 0.0666666, u = (-1, 0) on both sides of faces with normals
  (1, 0) and (-1, 0).  A uniform state's Godunov flux is its physical flux,
  written here in closed form: mass = rho u.n, momentum = rho u (u.n) + p n,
  energy = (E + p) u.n with E = p/(gamma - 1) + rho |u|^2 / 2; the kernel
  prints the mass flux, -1 and +1.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

TORO = {
    "gamma": 1.4,
    "source": "Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, Tables 4.1, 4.3",
    "tests": [
        {"name": "1 (Sod)", "left": [1.0, 0.0, 1.0], "right": [0.125, 0.0, 0.1],
         "p_star": 0.30313, "u_star": 0.92745, "rho_star_L": 0.42632, "rho_star_R": 0.26557},
        {"name": "2 (123)", "left": [1.0, -2.0, 0.4], "right": [1.0, 2.0, 0.4],
         "p_star": 0.00189, "u_star": 0.0, "rho_star_L": 0.02185, "rho_star_R": 0.02185},
        {"name": "3", "left": [1.0, 0.0, 1000.0], "right": [1.0, 0.0, 0.01],
         "p_star": 460.894, "u_star": 19.5975, "rho_star_L": 0.57506, "rho_star_R": 5.99924},
        {"name": "4", "left": [1.0, 0.0, 0.01], "right": [1.0, 0.0, 100.0],
         "p_star": 46.0950, "u_star": -6.19633, "rho_star_L": 5.99242, "rho_star_R": 0.57511},
        {"name": "5", "left": [5.99924, 19.5975, 460.894], "right": [5.99242, -6.19633, 46.0950],
         "p_star": 1691.64, "u_star": 8.68975, "rho_star_L": 14.2823, "rho_star_R": 31.0426},
    ],
}


def kat():
    g = 5.0 / 3.0  # Header.h:17-18
    rho, p, ux, uy = 1.0, 0.0666666, -1.0, 0.0
    E = p / (g - 1.0) + 0.5 * rho * (ux * ux + uy * uy)
    cases = []
    for n1, n2 in ((1.0, 0.0), (-1.0, 0.0)):
        un = ux * n1 + uy * n2
        cases.append({"normal": [n1, n2], "mass": rho * un,
                      "mom": [rho * ux * un + p * n1, rho * uy * un + p * n2],
                      "energy": (E + p) * un})
    return {"gamma": g, "state": {"rho": rho, "p": p, "u": [ux, uy]}, "cases": cases,
            "printed_mass_flux": [-1.0, 1.0], "source": "kernel.cu:3095-3107"}


if __name__ == "__main__":
    with open(os.path.join(HERE, "toro_table.json"), "w") as f:
        json.dump(TORO, f, indent=1)
    with open(os.path.join(HERE, "kat_reference_test_kernel.json"), "w") as f:
        json.dump(kat(), f, indent=1)
    print("wrote toro_table.json, kat_reference_test_kernel.json")
