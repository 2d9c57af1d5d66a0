"""Explicit Butcher tableaux (kernel/Base/TimeIntegration/*) and the rule
lookup of AllTimeIntegrators::getRule (AllTimeIntegrators.cpp:68-97).

Only the coefficients live here: the stage loop runs on the device.
"""
from dataclasses import dataclass
from typing import List, Optional


@dataclass(frozen=True)
class ButcherTableau:
    name: str
    order: int
    a: List[List[float]]  # strictly lower rows, row i has i entries
    b: List[float]
    c: List[float]
    tvd: bool = False
    error: Optional[List[float]] = None  # embedded error coefficients (b - b_low), if any

    @property
    def stages(self):
        return len(self.b)

    @property
    def has_error_estimate(self):
        return self.error is not None

    def a_matrix(self):
        s = self.stages
        m = [[0.0] * s for _ in range(s)]
        for i, row in enumerate(self.a):
            for j, v in enumerate(row):
                m[i][j] = float(v)
        return m


# ForwardEuler.cpp
FORWARD_EULER = ButcherTableau("ForwardEuler", 1, [[]], [1.0], [0.0])

# MidPoint.cpp
MIDPOINT = ButcherTableau("MidPoint", 2, [[], [0.5]], [0.0, 1.0], [0.0, 0.5])

# RK2TVD.cpp (Heun; embedded first-order estimate)
RK2_TVD = ButcherTableau("RK2TVD", 2, [[], [1.0]], [0.5, 0.5], [0.0, 1.0], tvd=True, error=[-0.5, 0.5])

# RK3TVD.cpp (Shu-Osher)
RK3_TVD = ButcherTableau("RK3TVD", 3, [[], [1.0], [0.25, 0.25]], [1 / 6, 1 / 6, 2 / 3], [0.0, 1.0, 0.5],
                         tvd=True)

# RK4_4: RK4Methods.cpp:45-63 (getRule(4, 4))
RK4 = ButcherTableau("RK4_4", 4, [[], [0.5], [0.0, 0.5], [0.0, 0.0, 1.0]],
                     [1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6], [0.0, 0.5, 0.5, 1.0])

# DormandPrince.cpp:45-80 (getRule(5, 7), the API default), with the
# published 4th-order embedded weights in the error coefficients
DORMAND_PRINCE = ButcherTableau(
    "DormandPrince", 5,
    [[], [1 / 5], [3 / 40, 9 / 40], [44 / 45, -56 / 15, 32 / 9],
     [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
     [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
     [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]],
    [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0],
    [0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0],
    error=[35 / 384 - 5179 / 57600, 0.0, 500 / 1113 - 7571 / 16695, 125 / 192 - 393 / 640,
           -2187 / 6784 + 92097 / 339200, 11 / 84 - 187 / 2100, -1 / 40])

# BogackiShampine.cpp (3(2) pair)
BOGACKI_SHAMPINE = ButcherTableau(
    "BogackiShampine", 3, [[], [0.5], [0.0, 0.75], [2 / 9, 1 / 3, 4 / 9]],
    [2 / 9, 1 / 3, 4 / 9, 0.0], [0.0, 0.5, 0.75, 1.0],
    error=[2 / 9 - 7 / 24, 1 / 3 - 1 / 4, 4 / 9 - 1 / 3, -1 / 8])

ALL = [FORWARD_EULER, MIDPOINT, RK2_TVD, RK3_TVD, RK4, DORMAND_PRINCE, BOGACKI_SHAMPINE]


def get_rule(order, stages, tvd=False):
    """AllTimeIntegrators::getRule: exact (order, stages, tvd) match first,
    then the same order and stages with the TVD flag relaxed; else an error."""
    for t in ALL:
        if t.order == order and t.stages == stages and t.tvd == tvd:
            return t
    for t in ALL:
        if t.order == order and t.stages == stages:
            return t
    raise ValueError(f"no Runge-Kutta method of order {order} with {stages} stages")
