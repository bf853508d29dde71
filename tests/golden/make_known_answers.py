"""Writes known_answers.json: the reference's own test constants and expected
results for this path, plus closed-form values, as data.

Sources (paths relative to the AndoSim tree):
  tests/test_barrier_derivatives.cpp:21-236   barrier V, V', V'' at g=0.005, gmax=0.01, k=1000
  tests/test_basic.cpp:70-81                  pin stiffness, m=0.1, w=(0.1,0,0), H=500 I
  tests/test_basic.cpp:193-219                point-triangle distance
  tests/test_basic.cpp:221-272                pin / wall gradient signs
  tests/test_basic.cpp:274-375                line-search alpha reductions
  tests/test_basic.cpp:377-419                9x9 PCG system
Closed forms:
  barrier: V = k/(2 gmax) (gmax-g)^3, V' = -3k/(2 gmax) (gmax-g)^2, V'' = 3k/gmax (gmax-g)
  right-isosceles face (0,0),(h,0),(h,h): K = k/h^2 [[1,-1],[-1,2]], k = A t mu, A = h^2/2
  free fall of an unconstrained triangle: v_after_step = (15/16) v_after_gravity
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    g, gmax, k = 0.005, 0.01, 1000.0
    d = gmax - g
    E, nu, t, h = 3e5, 0.3, 1e-3, 0.05
    mu = E / (2 * (1 + nu))
    kh2 = (h * h / 2) * t * mu / (h * h)
    data = {
        "barrier": {"g": g, "gmax": gmax, "k": k, "V": k / (2 * gmax) * d ** 3, "dV": -1.5 * k / gmax * d * d,
                    "d2V": 3 * k / gmax * d, "fd_rel_tol": 1e-3,
                    "domain_zero_at": [0.015, 0.01], "positive_at": -0.001},
        "pin_stiffness": {"mass": 0.1, "offset": [0.1, 0.0, 0.0], "H_diag": 500.0, "min_gap": 1e-6,
                          "expected": 500.0 + 0.1 / 0.01, "abs_tol": 1.0},
        "point_triangle": {"p": [0.25, 0.25, 0.5], "a": [0, 0, 0], "b": [1, 0, 0], "c": [0, 1, 0],
                           "distance": 0.5, "normal_z_min": 0.9},
        "pin_gradient": {"x": [1.0, 0.0, 0.0], "target": [0, 0, 0], "gmax": 2.0, "k": 1000.0, "sign_x": -1},
        "wall_gradient": {"x": [0.0, 0.0, 0.5], "normal": [0, 0, 1], "offset": 0.0, "gmax": 1.0, "k": 1000.0,
                          "sign_z": -1},
        "line_search_wall": {"verts": [[0, 0, 0.1], [1, 0, 0.1], [0.5, 1, 0.1]], "dir_z0": -0.15, "ext": 1.0,
                             "alpha_open_interval": [0.0, 1.0], "expected_alpha": 0.5},
        "line_search_pt": {"verts": [[0.5, 0.5, 0.5], [0, 0, 0], [1, 0, 0], [0.5, 1, 0]], "tri": [1, 2, 3],
                           "dir_z0": -1.0, "ext": 1.0, "alpha_below": 1.0},
        "pcg9": {"n": 9, "diag": 2.0, "off": [[0, 1, 0.5], [1, 0, 0.5], [3, 4, 0.3], [4, 3, 0.3]], "b": 1.0,
                 "tol": 1e-6, "max_iters": 100, "residual_l2_max": 1e-3},
        "elastic_face": {"h": h, "E": E, "nu": nu, "t": t, "k_over_h2": kh2,
                         "K_over_k_h2": [[1.0, -1.0], [-1.0, 2.0]]},
        "free_fall": {"velocity_ratio": 15.0 / 16.0},
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
