"""Explicit Butcher tableaux (kernel/Base/TimeIntegration/*).

Only the coefficients are needed: the stage loop itself runs on the device.
"""
from dataclasses import dataclass
from typing import List


@dataclass(frozen=True)
class ButcherTableau:
    name: str
    order: int
    a: List[List[float]]  # strictly lower rows, row i has i entries
    b: List[float]
    c: List[float]

    @property
    def stages(self):
        return len(self.b)

    def a_matrix(self):
        s = self.stages
        m = [[0.0] * s for _ in range(s)]
        for i, row in enumerate(self.a):
            for j, v in enumerate(row):
                m[i][j] = float(v)
        return m


# RK4_4: kernel/Base/TimeIntegration/RK4Methods.cpp:45-63 (getRule(4, 4))
RK4 = ButcherTableau("RK4_4", 4, [[], [0.5], [0.0, 0.5], [0.0, 0.0, 1.0]],
                     [1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6], [0.0, 0.5, 0.5, 1.0])

# DormandPrince: kernel/Base/TimeIntegration/DormandPrince.cpp:45-80 (getRule(5, 7), the API default)
DORMAND_PRINCE = ButcherTableau(
    "DormandPrince", 5,
    [[], [1 / 5], [3 / 40, 9 / 40], [44 / 45, -56 / 15, 32 / 9],
     [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
     [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
     [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]],
    [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0],
    [0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0])

# ForwardEuler: kernel/Base/TimeIntegration/ForwardEuler.cpp
FORWARD_EULER = ButcherTableau("ForwardEuler", 1, [[]], [1.0], [0.0])
