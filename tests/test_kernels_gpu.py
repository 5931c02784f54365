"""Kernel-level parity: the transform passes (qf_tile.hip) of both kernel
families, register-radix and quad-split, and of the 8192-element tiles of
the 10-bit strided passes, against numpy restatements of the same 2x2 tensor
maps, over F65521/Ext8 (libparcomm_amd.so) and KoalaBear/Ext4
(libparcomm_amd_kb.so). Each family is forced with the autotuner off
(QF_AUTOTUNE=0; QF_QTILE=0 keeps the register-radix kernels, the quad-split
ones take every geometry they support; QF_BTILE=0 keeps 10-bit strided passes
off the 8192-element tiles, which otherwise take them)."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _forced(family_name):
    saved = {k: os.environ.get(k) for k in ("QF_AUTOTUNE", "QF_QTILE", "QF_BTILE")}
    os.environ["QF_AUTOTUNE"] = "0"
    os.environ["QF_QTILE"] = "1" if family_name == "quad" else "0"
    os.environ["QF_BTILE"] = "1" if family_name == "big" else "0"
    return saved


def _restore(saved):
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.fixture(params=["radix", "quad"])
def family(request):
    saved = _forced(request.param)
    yield request.param
    _restore(saved)


@pytest.fixture
def big_tiles():
    saved = _forced("big")
    yield "big"
    _restore(saved)
# field -> (p, extension degree, X^degree wrap constant, coefficient dtype)
FIELDS = {0: (65521, 8, 17, np.uint16), 1: (2130706433, 4, 3, np.uint32)}
FIELD_IDS = {0: "f65521", 1: "koalabear"}


def e8_mul(a, b, field):
    """Rows of (..., lanes) canonical coefficients: a * b in F_p[X]/(X^lanes - w)."""
    P, lanes, wrap, _ = FIELDS[field]
    a = a.astype(np.int64)
    b = b.astype(np.int64)
    t = np.zeros(a.shape[:-1] + (2 * lanes - 1,), dtype=np.int64)
    for i in range(lanes):
        for j in range(lanes):
            t[..., i + j] = (t[..., i + j] + a[..., i] * b[..., j]) % P
    out = t[..., :lanes].copy()
    out[..., :lanes - 1] = (out[..., :lanes - 1] + wrap * t[..., lanes:]) % P
    return out % P


def transform(x, bits, wht, P):
    """x: (vectors, n, lanes); 2x2 maps on the given index bits."""
    y = x.astype(np.int64).copy()
    n = y.shape[1]
    idx = np.arange(n)
    for b in bits:
        lo = idx[((idx >> b) & 1) == 0]
        hi = lo | (1 << b)
        u, w = y[:, lo].copy(), y[:, hi].copy()
        if wht:
            y[:, lo] = (u + w) % P
            y[:, hi] = (u - w) % P
        else:
            y[:, hi] = (w - u) % P
    return y % P


def run(field, mode, x, alpha, L, s, t, vstride_in=None, vstride_out=None, xvec=False):
    import parcomm_amd as pa
    _, lanes, _, dt = FIELDS[field]
    vectors = x.shape[0]
    n = 1 << L
    vin = vstride_in or n
    vout = vstride_out or n
    host_in = np.zeros((vectors * vin, lanes), dtype=dt)
    for v in range(vectors):
        host_in[v * vin:v * vin + n] = x[v]
    d_in = pa.DeviceBuffer.from_host(host_in, field)
    d_o1 = pa.DeviceBuffer(vectors * vout * 16, field)
    d_o2 = pa.DeviceBuffer(vectors * vout * 16, field)
    d_a = pa.DeviceBuffer.from_host(alpha.astype(dt), field) if alpha is not None else None
    L_ = pa.lib(field)
    rc = L_.qf_test_tile_pass(mode, d_in.ptr, d_o1.ptr, d_o2.ptr, d_a.ptr if d_a else None, vectors, vin, vout,
                              L, s, t, int(xvec))
    assert rc == 0, L_.qf_last_error()
    o1 = d_o1.to_host(dt).reshape(vectors * vout, lanes)
    o2 = d_o2.to_host(dt).reshape(vectors * vout, lanes)
    take = lambda o: np.stack([o[v * vout:v * vout + n] for v in range(vectors)])
    return take(o1), take(o2)


GEOMS = [(32, 12, 0, 12), (2, 14, 12, 2), (1, 16, 12, 4), (16, 8, 0, 8), (1, 22, 12, 10), (1, 20, 12, 8),
         (32, 17, 12, 5), (2, 13, 0, 12), (1, 23, 22, 1)]


def field_values(rng, P, shape):
    """Uniform canonical coefficients with the extremes 0, 1, P - 1 planted."""
    x = rng.integers(0, P, size=shape, dtype=np.int64)
    flat = x.reshape(-1)
    flat[:3] = [0, 1, P - 1]
    flat[-3:] = [P - 1, P - 1, P - 2]
    return x


@pytest.mark.parametrize("field", [0, 1], ids=lambda f: FIELD_IDS[f])
@pytest.mark.parametrize("vectors,L,s,t", GEOMS)
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6])
def test_tile_pass_matches_numpy(vectors, L, s, t, mode, field, family):
    P, lanes, _, _ = FIELDS[field]
    rng = np.random.default_rng(1000 * L + 10 * mode + s + 7 * field)
    x = field_values(rng, P, (vectors, 1 << L, lanes))
    alpha = field_values(rng, P, (1 << L, lanes))
    print("tile pass", family, FIELD_IDS[field], "mode", mode, "vectors", vectors, "L", L, "bits", s, s + t,
          flush=True)
    o1, o2 = run(field, mode, x, alpha if mode in (3, 5, 6) else None, L, s, t,
                 vstride_in=(1 << L) + 64 if L < 20 else None)
    bits = list(range(s, s + t))
    if mode == 0:
        assert np.array_equal(o1, transform(x, bits, False, P))
    elif mode == 1:
        m = transform(x, bits, False, P)
        assert np.array_equal(o1, m)
        assert np.array_equal(o2, transform(m, bits, True, P))
    elif mode == 2:
        assert np.array_equal(o1, transform(x, bits, True, P))
    elif mode == 3:
        w = transform(x, bits, True, P)
        assert np.array_equal(o1, transform(e8_mul(w, alpha[None], field), bits, True, P))
    elif mode == 4:
        assert np.array_equal(o1, x)
        assert np.array_equal(o2, transform(x, bits, True, P))
    elif mode == 6:
        assert np.array_equal(o1, transform(e8_mul(x, alpha[None], field), bits, True, P))
    else:
        w = transform(x, bits, True, P)
        assert np.array_equal(o1, x)
        assert np.array_equal(o2, transform(e8_mul(w, alpha[None], field), bits, True, P))


# The 8192-element tiles: Mobius, Mobius then WHT, and WHT over 10 strided
# bits (8 contiguous positions x 2^10 rows), as in the 2^22-row commit's high
# row group, with several vectors, padded vector strides and a hi index.
BIG_GEOMS = [(1, 22, 12, 10), (4, 14, 3, 10), (2, 16, 5, 10), (1, 18, 7, 10)]


@pytest.mark.parametrize("field", [0, 1], ids=lambda f: FIELD_IDS[f])
@pytest.mark.parametrize("vectors,L,s,t", BIG_GEOMS)
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_big_tile_pass_matches_numpy(vectors, L, s, t, mode, field, big_tiles):
    P, lanes, _, _ = FIELDS[field]
    rng = np.random.default_rng(5000 + 10 * s + L + mode + 7 * field)
    x = field_values(rng, P, (vectors, 1 << L, lanes))
    print("big tile pass", FIELD_IDS[field], "mode", mode, "vectors", vectors, "L", L, "bits", s, s + t, flush=True)
    pad = (1 << L) + 4096 if L < 20 else None
    o1, o2 = run(field, mode, x, None, L, s, t, vstride_in=pad, vstride_out=pad)
    bits = list(range(s, s + t))
    if mode == 0:
        assert np.array_equal(o1, transform(x, bits, False, P))
    elif mode == 1:
        m = transform(x, bits, False, P)
        assert np.array_equal(o1, m)
        assert np.array_equal(o2, transform(m, bits, True, P))
    else:
        assert np.array_equal(o1, transform(x, bits, True, P))


# Geometries of the 2^23-row commit schedule: Mobius on row bits 7..14 and
# 15..22, the pre-alpha parity pass on bits 12..14 (a pass without exchanges)
# and the high parity groups.
PLAN23 = [(1, 23, 7, 8), (1, 23, 15, 8), (1, 23, 12, 3), (1, 23, 12, 6), (1, 23, 18, 5)]


@pytest.mark.parametrize("field", [0, 1], ids=lambda f: FIELD_IDS[f])
@pytest.mark.parametrize("vectors,L,s,t", PLAN23)
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_tile_pass_commit_plan_geometries(vectors, L, s, t, mode, field, family):
    P, lanes, _, _ = FIELDS[field]
    rng = np.random.default_rng(3000 + 10 * s + t + mode + 7 * field)
    x = field_values(rng, P, (vectors, 1 << L, lanes))
    print("plan tile pass", FIELD_IDS[field], "mode", mode, "L", L, "bits", s, s + t, flush=True)
    o1, o2 = run(field, mode, x, None, L, s, t)
    bits = list(range(s, s + t))
    if mode == 0:
        assert np.array_equal(o1, transform(x, bits, False, P))
    elif mode == 1:
        m = transform(x, bits, False, P)
        assert np.array_equal(o1, m)
        assert np.array_equal(o2, transform(m, bits, True, P))
    else:
        assert np.array_equal(o1, transform(x, bits, True, P))


def transform_vectors(x, nbits, wht, P):
    """2x2 maps on the low nbits of the vector index (axis 0)."""
    y = x.astype(np.int64).copy()
    idx = np.arange(y.shape[0])
    for b in range(nbits):
        lo = idx[((idx >> b) & 1) == 0]
        hi = lo | (1 << b)
        u, w = y[lo].copy(), y[hi].copy()
        if wht:
            y[lo] = (u + w) % P
            y[hi] = (u - w) % P
        else:
            y[hi] = (w - u) % P
    return y % P


@pytest.mark.parametrize("field", [0, 1], ids=lambda f: FIELD_IDS[f])
@pytest.mark.parametrize("L", [12, 15])
@pytest.mark.parametrize("mode", [0, 2])
def test_tile_pass_over_column_bits(L, mode, field, family):
    # 32 columns x 128 rows per tile: the first commit pass (column bits + row bits 0..6)
    P, lanes, _, _ = FIELDS[field]
    rng = np.random.default_rng(77 + L + mode + field)
    x = field_values(rng, P, (32, 1 << L, lanes))
    print("column-bit tile pass", FIELD_IDS[field], "mode", mode, "L", L, flush=True)
    o1, _ = run(field, mode, x, None, L, 0, 7, vstride_in=(1 << L) + 16, vstride_out=2 << L, xvec=True)
    wht = mode == 2
    assert np.array_equal(o1, transform_vectors(transform(x, range(7), wht, P), 5, wht, P))



@pytest.mark.parametrize("field", [0, 1], ids=lambda f: FIELD_IDS[f])
@pytest.mark.parametrize("mode", [0, 2])
def test_column_bit_pass_with_wide_vector_stride(mode, field, family):
    # The first commit pass of a 2^30-entry KoalaBear block: 32 codeword
    # columns 2^24 elements (256 MiB) apart, so a tile spans 8 GiB and the
    # per-slot offsets need 64 bits. Only the first 2^12 elements of every
    # column are transformed and copied back.
    import parcomm_amd as pa
    P, lanes, _, dt = FIELDS[field]
    L, vs = 12, 1 << 24
    rng = np.random.default_rng(91 + mode + field)
    x = field_values(rng, P, (32, 1 << L, lanes))
    d_in = pa.DeviceBuffer(32 * vs * 16, field)
    d_o1 = pa.DeviceBuffer(32 * vs * 16, field)
    for v in range(32):
        d_in.upload(x[v].astype(dt), v * vs * 16)
    print("wide-stride column-bit pass", FIELD_IDS[field], "mode", mode, flush=True)
    L_ = pa.lib(field)
    rc = L_.qf_test_tile_pass(mode, d_in.ptr, d_o1.ptr, None, None, 32, vs, vs, L, 0, 7, 1)
    assert rc == 0, L_.qf_last_error()
    o1 = np.stack([d_o1.to_host(dt, v * vs * 16, (1 << L) * lanes).reshape(1 << L, lanes) for v in range(32)])
    wht = mode == 2
    assert np.array_equal(o1, transform_vectors(transform(x, range(7), wht, P), 5, wht, P))
    d_in.close()
    d_o1.close()


@pytest.mark.parametrize("half", [None, 0, 1], ids=["whole", "half0", "half1"])
def test_device_copy_calibration_kernels(half):
    """The calibration copies: the whole range, or exactly one 64-byte half of
    every 128-byte line with the other half of dst left as it was."""
    import parcomm_amd as pa
    rng = np.random.default_rng(7)
    host = rng.integers(0, 1 << 32, size=(1 << 20) // 4, dtype=np.uint32)  # 1 MiB
    src = pa.DeviceBuffer.from_host(host)
    dst = pa.DeviceBuffer(host.nbytes)  # zeroed
    dst.copy_from(src, half=half)
    got = dst.to_host(np.uint32).reshape(-1, 32)  # 128-byte lines of 32 words
    want = host.reshape(-1, 32).copy()
    if half is not None:
        want[:, 16 * (1 - half):16 * (2 - half)] = 0
    assert np.array_equal(got, want)
    with pytest.raises(Exception):
        dst.copy_from(src, nbytes=48, half=0)  # not a whole number of lines
