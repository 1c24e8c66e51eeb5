"""Exact restatement of the split-precision product (test infrastructure only).

The library's split path (spla_amd/csrc/runtime/split_plan.cpp,
spla_amd/csrc/kernels/ozaki.hip) claims to compute, for alpha = 1, beta = 0,

    C_ij = conv(X_ij) * 2^-(ea_i + eb_j),   X = a^T b exactly (Python integers),

with a_il = rint(op(A)_il * 2^ea_i), ea_i = PA - 1 - floor(log2 max_l |op(A)_il|)
(componentwise for complex, conjugation folded into the imaginary part), the
same for b, and conv the conversion of the exact integer X to double through
its 64-bit limbs ((w2 * 2^128 + w1 * 2^64) + w0, each limb rounded to double).
Everything here is exact integer arithmetic, so the GPU result must match bit
for bit whenever the plan's moduli cover X. The plan restatement mirrors the
rules of split_plan.cpp so that both can be compared on the same statistics.
"""

from __future__ import annotations

import math
import operator

import numpy as np

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
          223, 217, 211, 199, 197, 193, 191, 181, 179, 173]


def ilogb(x: float) -> int:
    return math.frexp(x)[1] - 1


def fp64_bound(k: int, comps: int) -> float:
    return k * comps * 2.0 ** -53


def plan(maxA, sumsqA, maxB, sumsqB, k: int, comps: int, tol: float, max_moduli: int = 20):
    """Restatement of plan_split(): returns dict(ok, reason, PA, PB, moduli, bound, expA, expB)."""
    res = {"ok": False, "reason": 7, "PA": 0, "PB": 0, "moduli": 0, "bound": 0.0}
    for mx, sq in ((maxA, sumsqA), (maxB, sumsqB)):
        if not (np.all(np.isfinite(mx)) and np.all(np.isfinite(sq))):
            res["reason"] = 1
            return res
    if not (np.any(np.asarray(maxA) > 0) and np.any(np.asarray(maxB) > 0)):
        res["reason"] = 2
        return res
    nA, nB = float(np.sum(sumsqA)), float(np.sum(sumsqB))
    if not (nA > 0 and nB > 0 and math.isfinite(nA) and math.isfinite(nB)):
        res["reason"] = 5
        return res

    def scale_factor(mx):
        fl = [ilogb(v) for v in mx if v > 0]
        top = max(fl)
        acc = sum(math.ldexp(1.0, 2 * (f - top)) for f in fl)
        return math.ldexp(math.sqrt(comps * k * acc), top)

    SA = scale_factor(maxA) / math.sqrt(nA)
    SB = scale_factor(maxB) / math.sqrt(nB)
    half = tol / 2.02
    PA = max(1, math.ceil(math.log2(SA / half)))
    PB = max(1, math.ceil(math.log2(SB / half)))
    rhoA, rhoB = math.ldexp(SA, -PA), math.ldexp(SB, -PB)
    res.update(PA=PA, PB=PB, bound=rhoA + rhoB + rhoA * rhoB)
    if PA > 53 or PB > 53:
        res["reason"] = 3
        return res
    expA = [PA - 1 - ilogb(v) if v > 0 else 0 for v in maxA]
    expB = [PB - 1 - ilogb(v) if v > 0 else 0 for v in maxB]

    def row_norm_log2(mx, sq, e, P):
        cap = P + 0.5 * math.log2(comps * k)
        best = -math.inf
        for v, s, ee in zip(mx, sq, e):
            if not v > 0:
                continue
            est = cap
            if s > 0:
                a = ee + 0.5 * math.log2(s)
                b = math.log2(0.5 * math.sqrt(comps * k))
                hi, lo = max(a, b), min(a, b)
                est = min(cap, hi + math.log2(1.0 + 2.0 ** (lo - hi)))
            best = max(best, est)
        return best

    bits_needed = row_norm_log2(maxA, sumsqA, expA, PA) + row_norm_log2(maxB, sumsqB, expB, PB) + 2
    bits, n = 0.0, 0
    while n < len(MODULI) and bits < bits_needed:
        bits += math.log2(MODULI[n])
        n += 1
    res["moduli"] = n
    if bits < bits_needed or n > max_moduli:
        res["reason"] = 4
        return res
    res.update(ok=True, reason=0, expA=expA, expB=expB, bits_needed=bits_needed)
    return res


def crt_constants(count: int):
    """(W_t, M): W_t = (M/m_t) * ((M/m_t)^-1 mod m_t)."""
    M = 1
    for m in MODULI[:count]:
        M *= m
    W = []
    for m in MODULI[:count]:
        Mt = M // m
        W.append(Mt * pow(Mt % m, -1, m))
    return W, M


def limb_to_double(v: int) -> float:
    mag = abs(v)
    w0 = mag & ((1 << 64) - 1)
    w1 = (mag >> 64) & ((1 << 64) - 1)
    w2 = mag >> 128
    d = (math.ldexp(float(w2), 128) + math.ldexp(float(w1), 64)) + float(w0)
    return -d if v < 0 else d


def scaled_row(r: np.ndarray, P: int, cplx: bool):
    """Integers of one row of op(A) (or column of op(B)) at width P and its exponent:
    e = P - 1 - floor(log2 max |component|), entries rint(x 2^e) (round half to even)."""
    comps = np.concatenate([r.real, r.imag]) if cplx else r
    mx = float(np.max(np.abs(comps))) if comps.size else 0.0
    e = P - 1 - ilogb(mx) if mx > 0 else 0
    # ldexp is exact and |x 2^e| <= 2^P <= 2^53, so rint and int64 are exact
    re = np.rint(np.ldexp(np.asarray(r.real, dtype=np.float64), e)).astype(np.int64).tolist()
    im = (np.rint(np.ldexp(np.asarray(r.imag, dtype=np.float64), e)).astype(np.int64).tolist()
          if cplx else None)
    return re, im, e


def emulate_pairs(a_rows, b_cols, PA: int, PB: int):
    """Entries (i, j) of the split product (alpha = 1, beta = 0) for given rows a_rows[s] of
    op(A) and columns b_cols[s] of op(B), by the rules of emulate(); one value per pair."""
    out = []
    for r, c in zip(a_rows, b_cols):
        cplx = np.iscomplexobj(r)
        ar, ai, ea = scaled_row(r, PA, cplx)
        br, bi, eb = scaled_row(c, PB, cplx)
        sc = -(ea + eb)
        if cplx:
            re = sum(map(operator.mul, ar, br)) - sum(map(operator.mul, ai, bi))
            im = sum(map(operator.mul, ar, bi)) + sum(map(operator.mul, ai, br))
            out.append(complex(math.ldexp(limb_to_double(re), sc),
                               math.ldexp(limb_to_double(im), sc)))
        else:
            out.append(math.ldexp(limb_to_double(sum(map(operator.mul, ar, br))), sc))
    return out


def op_rows(X: np.ndarray, ld: int, op: int, rows: int, k: int) -> np.ndarray:
    """rows x k array of op(X) read from column-major storage with leading dimension ld."""
    if op == 0:
        M = X[: ld * k].reshape(k, ld).T[:rows, :]
    else:
        M = X[: ld * rows].reshape(rows, ld)[:, :k]
    return np.conj(M) if op == 2 else M


def emulate(opA, opB, m, n, k, A, lda, B, ldb, PA, PB):
    """Split product alpha = 1, beta = 0 as the library defines it; returns m x n array."""
    cplx = np.iscomplexobj(A)
    opa = op_rows(A, lda, opA, m, k)            # m x k
    if opB == 0:
        opbT = B[: ldb * n].reshape(n, ldb)[:, :k]  # column j of B as row j
    else:
        Bm = B[: ldb * k].reshape(k, ldb).T[:n, :]  # B stored n x k; op(B)^T row j = B[j, :]
        opbT = np.conj(Bm) if opB == 2 else Bm

    def scaled(rows, P):
        out, expo = [], []
        for r in rows:
            re, im, e = scaled_row(r, P, cplx)
            expo.append(e)
            out.append((re, im))
        return out, expo

    a, ea = scaled(opa, PA)
    b, eb = scaled(opbT, PB)
    C = np.empty((m, n), dtype=np.complex128 if cplx else np.float64)
    for i in range(m):
        ar, ai = a[i]
        for j in range(n):
            br, bi = b[j]
            sc = -(ea[i] + eb[j])
            if cplx:
                re = sum(x * y for x, y in zip(ar, br)) - sum(x * y for x, y in zip(ai, bi))
                im = sum(x * y for x, y in zip(ar, bi)) + sum(x * y for x, y in zip(ai, br))
                C[i, j] = complex(math.ldexp(limb_to_double(re), sc),
                                  math.ldexp(limb_to_double(im), sc))
            else:
                v = sum(x * y for x, y in zip(ar, br))
                C[i, j] = math.ldexp(limb_to_double(v), sc)
    return C
