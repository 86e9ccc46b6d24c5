"""Write the golden vectors under tests/golden/ from the reference tests' own fixtures.

The inputs restate the literal fixtures of the reference's unit tests; the expected outputs
come from numpy / scipy dense linear algebra in this container (SVD least squares of the
augmented damped system), never from the reference code. Run:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def column_scales(J, floor):
    return np.maximum(np.linalg.norm(J, axis=0), floor)


def damped_lstsq(J, r, D, lam):
    """argmin ||J d + r||^2 + lam ||D d||^2 via SVD least squares of [J; sqrt(lam) D]."""
    n = J.shape[1]
    A = np.vstack([J, np.sqrt(lam) * np.diag(D)])
    b = np.concatenate([-r, np.zeros(n)])
    x, *_ = np.linalg.lstsq(A, b, rcond=None)
    return x


def schur_fixture():
    """src/schur.rs:693-759: Schur::new(6, 2, 2), 12 coordinates, 5 two-row observations emitted
    in reverse block order, two unary rows on columns 0 and 8, floor 0.1."""
    obs = [(0, 6), (2, 6), (0, 8), (2, 8), (0, 6)]
    J = np.zeros((12, 12))
    r = np.zeros(12)
    emissions = []
    for i, (cam, pt) in enumerate(obs):
        a = np.array([[0.01, 2.0], [-0.03, 1.0 + i]])
        b = np.array([[10.0, -0.3], [20.0 + i, 1.0]])
        res = np.array([i - 2.0, 1.0])
        J[2 * i:2 * i + 2, cam:cam + 2] = a
        J[2 * i:2 * i + 2, pt:pt + 2] = b
        r[2 * i:2 * i + 2] = res
        emissions.append({"residual": res.tolist(), "blocks": [[pt, b.tolist()], [cam, a.tolist()]]})
    for i, col in enumerate([0, 8]):
        a = np.array([[0.5, -1.0]])
        J[10 + i, col:col + 2] = a
        r[10 + i] = 0.7
        emissions.append({"residual": [0.7], "blocks": [[col, a.tolist()]]})
    floor = 0.1
    D = column_scales(J, floor)
    lams = [0.001, 1.0, 100.0, 0.01]
    steps, preds = [], []
    for lam in lams:
        d = damped_lstsq(J, r, D, lam)
        steps.append(d.tolist())
        preds.append(float(0.5 * (r @ r - np.sum((J @ d + r) ** 2))))
    return {"retained": 6, "retained_width": 2, "eliminated_width": 2, "dimension": 12, "floor": floor,
            "emissions": emissions, "J": J.tolist(), "r": r.tolist(), "scales": D.tolist(), "lambdas": lams,
            "steps": steps, "predicted": preds}


def lsmr_products_fixture():
    """src/lsmr.rs:620-726: a 4x3 Jacobian emitted as two 2-row blocks at column 0, floor 0.1;
    forward and transpose damped products for lambda in {0, 0.01, 1e8}, scaled and unscaled."""
    J = np.array([[0.001, 2.0, 3.0], [0.002, -1.0, 4.0], [-0.001, 3.0, -2.0], [0.003, 1.0, 2.0]])
    D = column_scales(J, 0.1)
    x = np.array([[(i + k) - 1.25 for k in range(2)] for i in range(3)])
    y = np.array([[(2 * i + k) - 2.5 for k in range(2)] for i in range(7)])
    cases = []
    for scaled in [False, True]:
        for lam in [0.0, 0.01, 1e8]:
            diag = np.full(3, np.sqrt(lam)) if scaled else np.sqrt(lam) * D
            dense = np.zeros((7, 3))
            dense[:4] = J / D if scaled else J
            dense[4:] = np.diag(diag)
            cases.append({"scaled": scaled, "lambda": lam, "forward": (dense @ x).tolist(),
                          "transpose": (dense.T @ y).tolist()})
    return {"J": J.tolist(), "floor": 0.1, "scales": D.tolist(), "x": x.tolist(), "y": y.tolist(), "cases": cases}


def lsmr_svd_fixtures():
    """src/lsmr.rs:784-849 (mixed scales, floor 1e-3) and :857-899 (retries, floor 1.0)."""
    out = []
    J = np.array([[1e-4, 2.0, 0.0, 0.0], [2e-4, -1.0, 3e4, 0.0], [0.0, 1.0, -1e4, 0.0]])
    r = np.array([1.0, -2.0, 3.0])
    D = column_scales(J, 1e-3)
    lams = [0.1, 10.0, 1e-3, 1e6]
    out.append({"name": "mixed_scale", "J": J.tolist(), "r": r.tolist(), "floor": 1e-3, "scales": D.tolist(),
                "lambdas": lams, "steps": [damped_lstsq(J, r, D, l).tolist() for l in lams]})
    J = np.array([[1.0, 2.0, 0.0], [3.0, -1.0, 0.0], [0.0, 1.0, 0.0]])
    r = np.array([-3.0, 2.0, 1.0])
    D = column_scales(J, 1.0)
    lams = [1e-8, 0.01, 1.0, 100.0, 1e-4]
    out.append({"name": "retries", "J": J.tolist(), "r": r.tolist(), "floor": 1.0, "scales": D.tolist(),
                "lambdas": lams, "steps": [damped_lstsq(J, r, D, l).tolist() for l in lams]})
    return out


def main():
    data = {"schur_svd": schur_fixture(), "lsmr_products": lsmr_products_fixture(),
            "lsmr_svd": lsmr_svd_fixtures()}
    path = os.path.join(HERE, "linear_fixtures.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
