"""The parity tests' own helpers on hand-built cases (CPU): the contact
decision margin used to accept boundary flips."""
import numpy as np

from test_gpu_parity import decision_margin


def test_decision_margin_of_a_node_over_a_square_face():
    # unit square in z = 0, normal +z (x1 - x0) x (x2 - x0)
    x4 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0]])
    # 1e-7 m below the plane near the centre: the signed-distance end decides
    assert abs(decision_margin(x4, np.array([0.5, 0.4, -1e-7]), 0.3) - 1e-7) < 1e-12
    # 0.2 m below: 0.1 m from the max_distance end of the window
    assert abs(decision_margin(x4, np.array([0.5, 0.5, -0.2]), 0.3) - 0.1) < 1e-12
    # at the centre of the square, on the plane: 0 (the plane itself)
    assert decision_margin(x4, np.array([0.5, 0.5, 0.0]), 0.3) == 0.0
    # 0.05 m beyond the x = 1 edge and 0.15 m below: inside the 10 %
    # barycentric tolerance of the best triangle by 0.05, times the face
    # size sqrt(2); every other margin is 0.15 or more
    m = decision_margin(x4, np.array([1.05, 0.5, -0.15]), 0.3)
    assert abs(m - 0.05 * np.sqrt(2.0)) < 1e-12
