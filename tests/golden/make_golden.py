"""Writes gemm_golden.json: small GEMM cases with exactly-rounded results.

Every input is a double; the result alpha * op(A) * op(B) + beta * C is
evaluated in exact rational arithmetic (fractions.Fraction holds every double
exactly) and rounded once to the nearest double. These vectors pin the CPU
oracle (oracle/oracle.c) independently of any floating-point summation order.
The semantics are BLAS ?gemm's, which is what SPLA's CPU path calls
(src/gemm/gemm_host.cpp:55-72): op in {N, T, C}, column-major with leading
dimensions, beta == 0 never reads C, alpha == 0 or k == 0 never reads A / B.

Run: python tests/golden/make_golden.py  (deterministic; seed 2024)
"""

from __future__ import annotations

import json
import random
from fractions import Fraction
from pathlib import Path

OUT = Path(__file__).resolve().parent / "gemm_golden.json"


def rnd(r: random.Random) -> float:
    # mix of magnitudes and sign; full 53-bit mantissas
    return r.uniform(-1.0, 1.0) * 2.0 ** r.randint(-6, 6)


def hexf(x: float) -> str:
    return float.hex(x)


def opval(X, ld, op, row, col, cplx):
    if op == 0:
        v = X[row + col * ld]
    else:
        v = X[col + row * ld]
    if cplx and op == 2:
        v = (v[0], -v[1])
    return v


def exact_case(r, cplx, opA, opB, m, n, k, alpha, beta, nan_c=False):
    ra, ca = (m, k) if opA == 0 else (k, m)
    rb, cb = (k, n) if opB == 0 else (n, k)
    lda, ldb, ldc = ra + 1, rb + 2, m + 1

    def fill(count):
        if cplx:
            return [(rnd(r), rnd(r)) for _ in range(count)]
        return [rnd(r) for _ in range(count)]

    A = fill(lda * ca)
    B = fill(ldb * cb)
    C = fill(ldc * n)
    if nan_c:
        C = [(float("nan"), float("nan")) if cplx else float("nan") for _ in C]
    out = list(C)
    for j in range(n):
        for i in range(m):
            if cplx:
                sr = Fraction(0)
                si = Fraction(0)
                if not (alpha == (0.0, 0.0)):
                    for l in range(k):
                        a = opval(A, lda, opA, i, l, True)
                        b = opval(B, ldb, opB, l, j, True)
                        ar, ai, br, bi = (Fraction(x) for x in (*a, *b))
                        sr += ar * br - ai * bi
                        si += ar * bi + ai * br
                al_r, al_i = Fraction(alpha[0]), Fraction(alpha[1])
                re = al_r * sr - al_i * si
                im = al_r * si + al_i * sr
                if beta != (0.0, 0.0):
                    cr, ci = (Fraction(x) for x in C[i + j * ldc])
                    be_r, be_i = Fraction(beta[0]), Fraction(beta[1])
                    re += be_r * cr - be_i * ci
                    im += be_r * ci + be_i * cr
                out[i + j * ldc] = (float(re), float(im))
            else:
                s = Fraction(0)
                if alpha != 0.0:
                    for l in range(k):
                        s += Fraction(opval(A, lda, opA, i, l, False)) * Fraction(
                            opval(B, ldb, opB, l, j, False))
                v = Fraction(alpha) * s
                if beta != 0.0:
                    v += Fraction(beta) * Fraction(C[i + j * ldc])
                out[i + j * ldc] = float(v)

    def enc(seq):
        if cplx:
            return [[hexf(a), hexf(b)] for a, b in seq]
        return [hexf(a) for a in seq]

    return {
        "type": "z" if cplx else "d", "opA": opA, "opB": opB, "m": m, "n": n, "k": k,
        "lda": lda, "ldb": ldb, "ldc": ldc,
        "alpha": [hexf(alpha[0]), hexf(alpha[1])] if cplx else hexf(alpha),
        "beta": [hexf(beta[0]), hexf(beta[1])] if cplx else hexf(beta),
        "A": enc(A), "B": enc(B), "C": enc(C), "expected": enc(out),
    }


def main() -> None:
    r = random.Random(2024)
    cases = []
    sizes = [(1, 1, 1), (3, 2, 5), (5, 4, 7), (4, 6, 13), (7, 3, 2)]
    for (m, n, k) in sizes:
        for opA in (0, 1):
            for opB in (0, 1):
                cases.append(exact_case(r, False, opA, opB, m, n, k, rnd(r), rnd(r)))
        for opA in (0, 1, 2):
            for opB in (0, 1, 2):
                cases.append(exact_case(r, True, opA, opB, m, n, k, (rnd(r), rnd(r)),
                                        (rnd(r), rnd(r))))
    # BLAS special cases: beta == 0 with NaN in C, alpha == 0, k == 0
    cases.append(exact_case(r, True, 2, 0, 4, 3, 5, (1.5, -0.5), (0.0, 0.0), nan_c=True))
    cases.append(exact_case(r, False, 1, 0, 4, 3, 5, 0.75, 0.0, nan_c=True))
    cases.append(exact_case(r, True, 2, 0, 3, 3, 4, (0.0, 0.0), (2.0, 1.0)))
    cases.append(exact_case(r, False, 0, 0, 3, 2, 0, 1.0, -3.0))
    OUT.write_text(json.dumps({"generator": "tests/golden/make_golden.py", "cases": cases}))
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
