"""CPU pins of the mixed-precision oracle (oracle/mxp_oracle.py).

The rounding rules are pinned bit for bit against torch's own CPU
conversions (torch.bfloat16 and torch.float8_e4m3fn, round to nearest even),
the split against its defining identities, the tile scale against its
definition by brute force."""
import numpy as np
import pytest

from oracle import mxp_oracle as mo

torch = pytest.importorskip("torch")


def _bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


def _torch_round(x, dtype):
    return torch.from_numpy(np.asarray(x, dtype=np.float32)).to(dtype).float() \
        .numpy()


def test_bf16_round_matches_torch_random():
    rng = np.random.default_rng(1)
    for scale in (1e-30, 1e-5, 0.5, 1.0, 3.0e4, 1e30):
        x = (rng.standard_normal(20000) * scale).astype(np.float32)
        assert np.array_equal(_bits(mo.bf16_round(x)),
                              _bits(_torch_round(x, torch.bfloat16)))


def test_bf16_round_halfway_cases():
    # mantissas whose dropped 16 bits are exactly one half, just below and
    # just above, with even and odd kept parts
    base = np.array([0x3F800000, 0x3F810000, 0x40490000, 0xBF810000],
                    dtype=np.uint32)
    lows = np.array([0x8000, 0x7FFF, 0x8001, 0x0000, 0xFFFF], dtype=np.uint32)
    u = (base[:, None] | lows[None, :]).ravel()
    x = u.view(np.float32)
    assert np.array_equal(_bits(mo.bf16_round(x)),
                          _bits(_torch_round(x, torch.bfloat16)))


def test_e4m3_round_matches_torch():
    f8 = getattr(torch, "float8_e4m3fn", None)
    if f8 is None:
        pytest.skip("torch without float8_e4m3fn")
    rng = np.random.default_rng(2)
    vals = [rng.uniform(-448, 448, 20000),
            rng.uniform(-1, 1, 20000),
            rng.uniform(-2.0 ** -5, 2.0 ** -5, 20000),   # subnormal region
            rng.uniform(-2.0 ** -9, 2.0 ** -9, 5000)]     # rounds to 0 / 2^-9
    # every e4m3 value and the midpoints between neighbours (ties)
    grid = np.array(sorted({float(v) for v in
                            _torch_round(np.linspace(0, 448, 200000), f8)}))
    mids = (grid[1:] + grid[:-1]) / 2
    vals += [grid, -grid, mids, -mids]
    x = np.concatenate(vals).astype(np.float32)
    x = x[np.abs(x) <= 448]
    assert np.array_equal(mo.e4m3_round(x), _torch_round(x, f8))


@pytest.mark.parametrize("terms", [1, 2, 3])
def test_bf16_split_identities(terms):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(50000) * rng.choice([1e-3, 1, 1e3], 50000)) \
        .astype(np.float32)
    parts = mo.bf16_split(x, terms)
    r = x.astype(np.float64)
    for h in parts:
        # each term is a BF16 number and each residual is exact
        assert np.array_equal(mo.bf16_round(h), h)
        r = r - h.astype(np.float64)
    rel = np.abs(r) / np.abs(x.astype(np.float64))
    # unit roundoff of 8, 16 and 24 significant bits
    bound = {1: 2.0 ** -8, 2: 2.0 ** -16, 3: 2.0 ** -24}[terms]
    assert rel.max() <= bound


def test_fp8_tile_exponent_rule():
    rng = np.random.default_rng(4)
    for amax in np.concatenate([rng.uniform(1e-8, 1e4, 2000),
                                [448.0, 224.0, 1.0, 2.0 ** -20, 3.5]]):
        e = mo.fp8_tile_exp(float(amax))
        assert np.ldexp(amax, e) <= 448.0 < np.ldexp(amax, e + 1)
    assert mo.fp8_tile_exp(0.0) == 0


def test_update_product_accuracy_order():
    """The emulated products approach L@U as the precision rises."""
    rng = np.random.default_rng(5)
    L = rng.uniform(-1, 1, (300, 200)).astype(np.float32)
    U = rng.uniform(-0.5, 0.5, (200, 260)).astype(np.float32)
    exact = L.astype(np.float64) @ U.astype(np.float64)
    absp = mo.update_abs_product(0, 1, L, U)
    errs = {}
    for upd, terms in [(1, 1), (1, 2), (1, 3), (2, 1)]:
        errs[(upd, terms)] = np.max(np.abs(mo.update_product(upd, terms, L, U)
                                           - exact) / absp)
    assert errs[(1, 3)] < 1e-7 < errs[(1, 2)] < 1e-4 < errs[(1, 1)] < 1e-2
    assert errs[(2, 1)] < 0.1
    assert np.array_equal(mo.update_product(0, 1, L, U), exact)


def test_fp8_tile_scales_exact_on_scaled_e4m3_tiles():
    """update_product(FP8) equals L @ U exactly when every tile is an e4m3
    matrix times its own power of two, and the exponent grids come out as
    built; a transposed or shifted exponent grid (planted here in the
    oracle's own formula) changes the product."""
    L, U, sL, sU = mo.scaled_e4m3_operands()
    exact = L.astype(np.float64) @ U.astype(np.float64)
    ql, el = mo.fp8_quantize(L, True)
    qu, eu = mo.fp8_quantize(U, False)
    assert np.array_equal(el, sL) and np.array_equal(eu, sU)
    assert np.array_equal(np.ldexp(ql.astype(np.float64),
                                   -np.repeat(np.repeat(el, 128, 0), 128, 1)
                                   [:L.shape[0], :L.shape[1]]), L)
    assert np.array_equal(mo.update_product(2, 1, L, U), exact)
    # the same product with a mis-indexed exponent grid
    K = L.shape[1]

    def with_grids(gl, gu):
        out = np.zeros(exact.shape)
        for kt in range(K // 128):
            ks = slice(kt * 128, (kt + 1) * 128)
            s_l = np.repeat(np.ldexp(1.0, -gl[:, kt]), 128)[:L.shape[0]]
            s_u = np.repeat(np.ldexp(1.0, -gu[kt, :]), 128)[:U.shape[1]]
            out += (ql[:, ks].astype(np.float64) * s_l[:, None]) @ \
                (qu[ks, :].astype(np.float64) * s_u[None, :])
        return out

    assert np.array_equal(with_grids(el, eu), exact)
    for gl, gu in [(el.T, eu), (el, eu.T), (np.roll(el, 1, 0), eu),
                   (el, np.roll(eu, 1, 1)), (np.roll(el, 1, 1), eu)]:
        assert not np.array_equal(with_grids(gl, gu), exact)
