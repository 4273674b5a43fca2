"""Writes mra_known_answers.json: encode/decode known answers of the
multiwavelet (MW) and Haar (HW) filter banks, computed from the closed forms
of the filter entries in 50-digit decimal arithmetic, independent of any
implementation.

Inputs are the reference's own known-answer children (1, 2, 3, 4 for Haar;
1..12 for MW), source/unittests/unit_tests_mra.cu:7-66. The filter entries
are the closed forms of the literals in source/mra/Filters.h:202-397:
sqrt(3)/4, 1/sqrt(14), sqrt(3/14), 1/sqrt(2), sqrt(3/112), 3/sqrt(112),
1/sqrt(7), 1/2, 1/4. Filter k of a family pairs with child PERM[k] = 0,2,1,3
(source/mra/encode_scale.cuh:16-19); encode is half the filtered sum.

    python tests/golden/make_mra_known_answers.py
"""
import json
import os
from decimal import Decimal, getcontext

getcontext().prec = 50
D = Decimal
S3_4 = D(3).sqrt() / 4
R14 = 1 / D(14).sqrt()
S3_14 = (D(3) / 14).sqrt()
R2 = 1 / D(2).sqrt()
S3_112 = (D(3) / 112).sqrt()
T112 = 3 / D(112).sqrt()
R7 = 1 / D(7).sqrt()
H, Q, Z = D("0.5"), D("0.25"), D(0)


def closed_form_filters():
    """[family][k] 3x3 closed forms, family order HH, GA, GB, GC."""
    sx = [-1, -1, 1, 1]     # filter k -> child PERM[k] x side (0 -> -, 1 -> +)
    sy = [-1, 1, -1, 1]
    HH = [[[H, Z, Z], [sx[k] * S3_4, Q, Z], [sy[k] * S3_4, Z, Q]] for k in range(4)]
    zero = [[Z] * 3 for _ in range(3)]
    GA = [[[-R14, -S3_14, -S3_14], [Z, Z, -R2], [Z, -R2, Z]], zero, zero,
          [[R14, -S3_14, -S3_14], [Z, Z, R2], [Z, R2, Z]]]
    GB = [zero, [[-R14, -S3_14, S3_14], [Z, Z, -R2], [Z, -R2, Z]],
          [[R14, -S3_14, S3_14], [Z, Z, R2], [Z, R2, Z]], zero]
    GC = [[[H, Z, Z], [-S3_112, -T112, R7], [-S3_112, R7, -T112]],
          [[-H, Z, Z], [-S3_112, -T112, -R7], [S3_112, -R7, -T112]],
          [[-H, Z, Z], [S3_112, -T112, -R7], [-S3_112, -R7, -T112]],
          [[H, Z, Z], [S3_112, -T112, R7], [S3_112, R7, -T112]]]
    return [HH, GA, GB, GC]


PERM = [0, 2, 1, 3]


def encode_mw(children):
    """children[c][m], c = child (bit0 x, bit1 y), m = 0,1x,1y -> 12 outputs
    ordered scale(0,1x,1y), alpha(0,1x,1y), beta(...), gamma(...)."""
    F = closed_form_filters()
    out = []
    for fam in range(4):
        for i in range(3):
            s = sum(F[fam][k][i][j] * children[PERM[k]][j] for k in range(4) for j in range(3))
            out.append(s / 2)
    return out


def encode_hw(c):
    r = R2
    s = H * (r * (r * c[0] + r * c[2]) + r * (r * c[1] + r * c[3]))
    a = H * (r * (r * c[0] - r * c[2]) + r * (r * c[1] - r * c[3]))
    b = H * (r * (r * c[0] + r * c[2]) - r * (r * c[1] + r * c[3]))
    g = H * (r * (r * c[0] - r * c[2]) - r * (r * c[1] - r * c[3]))
    return [s, a, b, g]


def main():
    mw_children = [[D(1 + 3 * c + m) for m in range(3)] for c in range(4)]
    hw_children = [D(1), D(2), D(3), D(4)]
    F = closed_form_filters()
    out = {
        "source": "closed forms of source/mra/Filters.h:202-397; inputs of source/unittests/unit_tests_mra.cu:7-66",
        "filters": [[[[str(F[f][k][i][j]) for j in range(3)] for i in range(3)] for k in range(4)] for f in range(4)],
        "mw_children": [[str(v) for v in row] for row in mw_children],
        "mw_encoded": [str(v) for v in encode_mw(mw_children)],
        "hw_children": [str(v) for v in hw_children],
        "hw_encoded": [str(v) for v in encode_hw(hw_children)],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mra_known_answers.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
