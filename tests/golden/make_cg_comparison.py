"""Writes cg_comparison.json: the 8x8 SPD system of the reference tests
(Tests/BiCGSTAB/CgComparisonData.cs:8-46) and the hand-computed values of the
workbook "CG сравнение.xlsm" (sheets LDLT_Result, расчет, точное решение; cell
data only).  Data only: matrix, right-hand side, exact solution, natural-order
LDL^T factor and diagonal, and plain-CG iterates.
"""
import json
import os

d = 741.7582417582419
a, b, c, e = -192.30769230769232, 357.14285714285717, -549.4505494505495, -164.83516483516485
matrix = [
    [d, 0, 0, 0, 0, 0, 0, 0],
    [0, d, 0, 0, 0, 0, 0, 0],
    [0, 0, d, 0, a, a, 0, b],
    [0, 0, 0, d, 0, 0, 0, 0],
    [0, 0, a, 0, d, b, c, a],
    [0, 0, a, 0, b, d, e, a],
    [0, 0, 0, 0, c, e, d, 0],
    [0, 0, b, 0, a, a, 0, d],
]
golden = {
    "source": "Tests/BiCGSTAB/CgComparisonData.cs:8-46; CG сравнение.xlsm sheets LDLT_Result, расчет, точное решение",
    "matrix": matrix,
    "rhs": [0.0, 0.0, 0.0, 0.0, 200.0, 0.0, 0.0, -1000.0],
    "exact_solution": [0.0, 0.0, 0.8379999999999997, 0.0, 0.31800000000000006, -0.358, 0.15600000000000014,
                       -1.7619999999999996],
    # natural-order LDL^T (LDLT_Result!A13:A20 and A2:H9)
    "ldlt_diagonal": [741.75824175824187, 741.75824175824187, 741.75824175824187, 741.75824175824187,
                      691.90069190069198, 555.42986425339382, 294.1406861977128, 532.62316910785626],
    "ldlt_lower": [[4, 2, -0.25925925925925924], [5, 2, -0.25925925925925924], [5, 4, 0.44411764705882345],
                   [6, 4, -0.79411764705882348], [6, 5, 0.14256619144602839], [7, 2, 0.48148148148148145],
                   [7, 4, -0.14411764705882352], [7, 5, -9.9796334012219962e-2], [7, 6, -0.24234354194407451]],
    # plain CG from x0 = 0 (sheet расчет)
    "cg_r0_dot": 1040000.0,
    "cg_p0_A_p0": 848351648.35164857,
    "cg_alpha0": 1.2259067357512951e-3,
    "cg_beta0": 0.39147359660662018,
    "cg_r0_norm": 1019.803902718557,
    "cg_r1_norm": 638.06938531078652,
    "cg_nonzero_components": [2, 4, 5, 6, 7],
    "cg_x_after_3": [0.80037154025645929, 0.20831569913302622, -0.37345333761090899, 3.3625935250457484e-2,
                     -1.7787989891728888],
    "cg_x_after_4": [0.83787140169131924, 0.23866523493131084, -0.32167702406718951, 9.8796140654675194e-2,
                     -1.7762388473545936],
    "ainv_3_3": 1.8774999999999994e-3,
}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cg_comparison.json"), "w") as f:
    json.dump(golden, f, indent=1)
