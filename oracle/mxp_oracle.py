"""TEST INFRASTRUCTURE ONLY -- CPU restatement of the mixed-precision update.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import
this module; the library never calls it.

The reference (rocHPL) is FP64 only: this restates the rounding rules of the
build's low-precision trailing update (rochpl_amd/csrc/k_mxp_gemm.hip), so
that a test can compute in FP64 exactly the products the MFMA units are given:

  * BF16 split: x ~ h0 + h1 + h2, h0 = bf16(x), h1 = bf16(x - h0),
    h2 = bf16(x - h0 - h1), bf16 = round to nearest even to 8 significant
    bits (the FP32 bit pattern rounded at bit 16); the product keeps the
    term pairs (i, j) with i + j < terms.
  * FP8: 128 x 128 tiles (row tile x K stage for L, K stage x column tile for
    U), scale 2^e with the largest e such that max|tile| 2^e <= 448, values
    rounded to OCP e4m3 (bias 7, 3 mantissa bits, subnormals 2^-9 apart,
    round to nearest even), the product of two tiles multiplied by
    2^-(eL + eU).

Pins (tests/test_mxp_oracle.py): the BF16 and e4m3 rounding against torch's
own CPU conversions (torch.bfloat16, torch.float8_e4m3fn) over random and
edge-case values, bit for bit; the split identities (terms are exact
differences); the scale rule against its definition by brute force; the
tile-scale placement by operands whose FP8 product is exact
(scaled_e4m3_operands: update_product must equal L @ U, and a transposed or
shifted exponent grid must not).
"""
from __future__ import annotations

import numpy as np

E4M3_MAX = 448.0


def bf16_round(x) -> np.ndarray:
    """float32 -> float32 values representable in BF16 (RNE)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(x.shape)


def bf16_split(x, terms: int) -> list[np.ndarray]:
    """The `terms` BF16 numbers whose sum approximates x (float32)."""
    r = np.asarray(x, dtype=np.float32).copy()
    out = []
    for _ in range(terms):
        h = bf16_round(r)
        out.append(h)
        r = (r - h).astype(np.float32)  # exact in FP32
    return out


def e4m3_round(x) -> np.ndarray:
    """float32 -> float32 values of OCP FP8 e4m3 (RNE), |x| <= 448."""
    x = np.asarray(x, dtype=np.float32)
    a = np.abs(x).astype(np.float64)
    sign = np.sign(x).astype(np.float64)
    out = np.zeros_like(a)
    # normal range: exponent e >= -6, 3 mantissa bits -> spacing 2^(e-3)
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    e = np.where(a > 0, e, -7)
    # guard log2 rounding
    e = np.where(2.0 ** e > a, e - 1, e)
    e = np.where(2.0 ** (e + 1) <= a, e + 1, e)
    normal = e >= -6
    spacing = np.where(normal, 2.0 ** (e - 3), 2.0 ** -9)
    q = a / spacing
    qr = np.round(q)  # numpy rounds half to even
    out = qr * spacing
    return (sign * out).astype(np.float32)


def fp8_tile_exp(amax: float) -> int:
    """Largest e with amax * 2^e <= 448 (0 for an all-zero tile)."""
    if not amax > 0:
        return 0
    e = int(np.floor(np.log2(E4M3_MAX / amax)))
    while np.ldexp(amax, e) > E4M3_MAX:
        e -= 1
    while np.ldexp(amax, e + 1) <= E4M3_MAX:
        e += 1
    return max(-100, min(100, e))


def fp8_quantize(M, tile_rows: bool):
    """Per-128x128-tile scaled e4m3 values of M (float32, rows x K for L with
    tile_rows=True, K x cols for U with tile_rows=False).  Returns
    (values as float32 e4m3 numbers, exponent grid [row tile, K tile] or
    [K tile, col tile])."""
    M = np.asarray(M, dtype=np.float32)
    R, C = M.shape
    q = np.zeros_like(M)
    ex = np.zeros(((R + 127) // 128, (C + 127) // 128), dtype=np.int64)
    for bi in range(ex.shape[0]):
        for bj in range(ex.shape[1]):
            blk = M[bi * 128:(bi + 1) * 128, bj * 128:(bj + 1) * 128]
            e = fp8_tile_exp(float(np.max(np.abs(blk))) if blk.size else 0.0)
            ex[bi, bj] = e
            s = np.float32(np.ldexp(1.0, e))
            q[bi * 128:(bi + 1) * 128, bj * 128:(bj + 1) * 128] = \
                e4m3_round((blk * s).astype(np.float32))
    return q, ex


def update_product(update: int, terms: int, L, U) -> np.ndarray:
    """The exact (FP64) value of the product the low-precision update adds,
    before FP32 accumulation rounding: C -= result.
    update: 0 FP32, 1 BF16 (terms 1..3), 2 FP8."""
    L = np.asarray(L, dtype=np.float32)
    U = np.asarray(U, dtype=np.float32)
    if update == 0:
        return L.astype(np.float64) @ U.astype(np.float64)
    if update == 1:
        ls, us = bf16_split(L, terms), bf16_split(U, terms)
        out = np.zeros((L.shape[0], U.shape[1]))
        for i in range(terms):
            for j in range(terms - i):
                out += ls[i].astype(np.float64) @ us[j].astype(np.float64)
        return out
    if update == 2:
        ql, el = fp8_quantize(L, True)
        qu, eu = fp8_quantize(U, False)
        K = L.shape[1]
        out = np.zeros((L.shape[0], U.shape[1]))
        for kt in range((K + 127) // 128):
            ks = slice(kt * 128, (kt + 1) * 128)
            # scale the row/column tiles of this K stage back
            sl = np.repeat(np.ldexp(1.0, -el[:, kt]), 128)[:L.shape[0]]
            su = np.repeat(np.ldexp(1.0, -eu[kt, :]), 128)[:U.shape[1]]
            out += (ql[:, ks].astype(np.float64) * sl[:, None]) @ \
                   (qu[ks, :].astype(np.float64) * su[None, :])
        return out
    raise ValueError(update)


def update_abs_product(update: int, terms: int, L, U) -> np.ndarray:
    """sum_k |l_ik| |u_kj| of the rounded operands (for error bounds)."""
    L = np.abs(np.asarray(L, dtype=np.float64))
    U = np.abs(np.asarray(U, dtype=np.float64))
    return L @ U


def scaled_e4m3_operands(seed=6):
    """L (300 x 384) and U (384 x 260) built tile by tile (128 x 128) from
    small e4m3 integers times a power of two that differs for every tile:
    L tile (row tile bi, K tile kt) carries 2^-(bi + 3 kt), U tile (K tile
    kt, column tile bj) 2^-(3 bj - kt).  Each tile holds one entry of
    magnitude 448 (so its scale exponent is exactly that tile's power) and
    otherwise integers in [-15, 15]: quantisation is exact, and every
    product, partial sum and output of the FP8 update is exactly
    representable in FP32 (for one output element the K tiles' powers span
    2^4 and each 128-deep integer sum stays below 2^18)."""
    rng = np.random.default_rng(seed)
    R, K, C = 300, 384, 260
    sL = np.array([[bi + 3 * kt for kt in range(3)] for bi in range(3)])
    sU = np.array([[3 * bj - kt for bj in range(3)] for kt in range(3)])

    def build(rows, cols, s_of):
        M = np.zeros((rows, cols), np.float64)
        for ti in range((rows + 127) // 128):
            for tj in range((cols + 127) // 128):
                r0, c0 = ti * 128, tj * 128
                r1, c1 = min(rows, r0 + 128), min(cols, c0 + 128)
                q = rng.integers(-15, 16, (r1 - r0, c1 - c0)).astype(np.float64)
                q[rng.integers(r1 - r0), rng.integers(c1 - c0)] = \
                    448.0 * rng.choice([-1.0, 1.0])
                M[r0:r1, c0:c1] = np.ldexp(q, -int(s_of(ti, tj)))
        return M.astype(np.float32)

    L = build(R, K, lambda bi, kt: sL[bi, kt])
    U = build(K, C, lambda kt, bj: sU[kt, bj])
    return L, U, sL, sU
