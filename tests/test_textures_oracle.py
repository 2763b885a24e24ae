"""The oracle's material-texture path (oracle/frame.c: oracle_texel_coord, oracle_linear_table,
oracle_interpret) pinned against what the mathematics fixes: texel centres map to their texel,
the lookup wraps with period 1 in both directions, the linearisation table is the correctly
rounded pow(c / 255, 2.2) (checked with 60-digit decimal arithmetic), and a hand-built
triangle + texture reads the texel, channel and row the coordinates name. CPU only."""
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import oracle as O
from phantasyenginetestbed_amd import phx

DIMS = [1, 2, 3, 5, 7, 8, 64, 100, 257, 1000, 4096, 65535, 65536]


@pytest.mark.parametrize("dim", DIMS)
def test_texel_centres_map_to_their_texel(dim):
    idx = np.unique(np.linspace(0, dim - 1, min(dim, 300)).astype(np.int64))
    for i in idx:
        c = float(np.float32((i + 0.5) / dim))
        assert O.texel_coord(c, dim) == i, (dim, i)


@pytest.mark.parametrize("dim", DIMS)
def test_lookup_wraps_with_period_one(dim):
    # Dyadic coordinates keep c - 1, c + 1 and c + 3 exact in float.
    rng = np.random.default_rng(dim)
    for c in (rng.integers(0, 1 << 12, 200) / 4096.0):
        base = O.texel_coord(float(c), dim)
        assert 0 <= base < dim
        for shift in (-3.0, -1.0, 1.0, 3.0):
            assert O.texel_coord(float(c) + shift, dim) == base, (dim, c, shift)


@pytest.mark.parametrize("dim", [1, 2, 8, 64, 4096, 65536])
def test_power_of_two_dims_floor_exactly(dim):
    # For a power-of-two dim, x = c * dim is exact and the texel is floor(x) for c in [0, 1),
    # except that the "+ dim" step rounds x + dim to a multiple of dim * 2^-23: when x lies within
    # half that step below the next integer, the lookup lands on the next texel (mod dim).
    rng = np.random.default_rng(7)
    c = np.concatenate([rng.random(2000).astype(np.float32),
                        np.float32(1.0) - np.float32(2.0) ** -np.arange(1, 25, dtype=np.float32),
                        (np.arange(1, 50, dtype=np.float32) - np.float32(2.0 ** -12)) / 50,
                        # m / dim - 2^-25 with m / dim in [1/8, 1/4): exact, and x sits
                        # dim * 2^-25 below the integer m, so "+ dim" carries it to texel m.
                        (np.arange(dim // 8, dim // 4) / dim - 2.0 ** -25).astype(np.float32)])
    c = c[(c >= 0) & (c < 1)]
    x = c.astype(np.float64) * dim
    half = dim * 2.0 ** -24
    gap = np.floor(x) + 1 - x
    want = np.floor(x).astype(np.int64)
    want = np.where(gap < half, (want + 1) % dim, want)
    got = np.array([O.texel_coord(float(v), dim) for v in c])
    ok = (got == want) | (gap == half)  # an exact tie goes to the even neighbour
    assert ok.all(), (c[~ok][:5], got[~ok][:5], want[~ok][:5])
    if dim >= 8:
        assert (gap < half).any()  # the carry case is exercised


def test_large_and_non_finite_coordinates():
    rng = np.random.default_rng(3)
    for dim in (3, 100, 1000, 65535):
        for c in rng.uniform(-1e6, 1e6, 500).astype(np.float32):
            assert 0 <= O.texel_coord(float(c), dim) < dim
    for c in (float("nan"), float("inf"), -float("inf")):
        assert O.texel_coord(c, 17) == 0
    # -0.0 and tiny negatives wrap to the last texel or to 0, never below 0.
    assert O.texel_coord(-0.0, 10) == 0
    assert O.texel_coord(-1e-30, 10) == 0  # -1e-29 + 10 rounds to 10, which wraps to 0
    assert O.texel_coord(-0.05, 10) == 9


def _f32_exact(x: np.float32) -> Decimal:
    return Decimal(float(x))


def test_linear_table_is_correctly_rounded_pow():
    getcontext().prec = 60
    table = O.linear_table()
    e = Decimal(float(np.float32(2.2)))
    assert table[0] == 0.0 and table[255] == 1.0
    for c in range(1, 256):
        x = np.float32(c) / np.float32(255.0)
        y = (e * _f32_exact(x).ln()).exp()
        got = table[c]
        err = abs(_f32_exact(got) - y)
        for nb in (np.nextafter(got, np.float32(0)), np.nextafter(got, np.float32(2))):
            assert err <= abs(_f32_exact(nb) - y), (c, got, nb)


def _one_triangle(uv0, uv1, uv2, mat=0):
    data = np.zeros(1, phx.TRIDATA)
    data["n"] = (0.0, 0.0, 1.0)
    data["uv"] = (uv0, uv1, uv2)
    data["mat"] = mat
    return data


def _records(datas, mats, hits_uv, textures):
    n = len(hits_uv)
    rays = np.zeros(n, phx.RAY)
    rays["d"] = (0.0, 0.0, -1.0)
    rays["tmax"] = 1e30
    hits = np.zeros(n, phx.HIT)
    hits["tri"] = 0
    hits["t"] = 1.0
    hits["u"] = [u for u, _ in hits_uv]
    hits["v"] = [v for _, v in hits_uv]
    return O.interpret(datas, mats, rays, hits, textures)


def test_interpret_reads_the_named_texel_channel_and_row():
    # uv = (u, v) at barycentrics (u, v): uv0 = (0,0), uv1 = (1,0), uv2 = (0,1).
    datas = _one_triangle((0, 0), (1, 0), (0, 1))
    W_, H_ = 4, 3
    rgba = np.zeros((H_, W_, 4), np.uint8)
    for y in range(H_):
        for x in range(W_):
            rgba[y, x] = (10 * x + 1, 10 * y + 2, 7 * x + 5 * y + 3, 40 * y + x)
    gray = (np.arange(H_ * W_, dtype=np.uint8).reshape(H_, W_) * 20 + 5)
    # Same grey image stored with a padded pitch: the padding (0xEE) must never be read.
    padded = np.full((H_, W_ * 2 + 6), 0xEE, np.uint8)
    padded[:, 0:W_ * 2:2] = gray
    padded[:, 1:W_ * 2:2] = 200
    mats = np.zeros(1, phx.MATERIAL)
    mats["itex"] = (0, 1, 2, -1)
    mats["albedo"] = (0.5, 0.5, 0.5, 0.25)
    mats["f2"] = (0.3, 0.7, 0, 0)
    lin = O.linear_table()
    texs = [rgba, gray, (padded, W_, 2)]
    pts = [(x + 0.5) / W_ for x in range(W_)]
    uvs = [(u, (y + 0.5) / H_) for y in range(H_) for u in pts]
    rec = _records(datas, mats, uvs, texs)
    for k, (u, v) in enumerate(uvs):
        x, y = int(u * W_), int(v * H_)
        r, g, b, a = rgba[y, x]
        assert tuple(rec["albedo"][k]) == (lin[r], lin[g], lin[b]), (x, y)
        assert rec["alpha"][k] == np.float32(a) / np.float32(255)
        assert rec["roughness"][k] == np.float32(gray[y, x]) / np.float32(255)
        assert rec["metallic"][k] == np.float32(gray[y, x]) / np.float32(255)


def test_interpret_fallbacks_to_constants():
    datas = _one_triangle((0, 0), (1, 0), (0, 1))
    gray = np.full((2, 2), 77, np.uint8)
    two = np.full((2, 2, 2), 90, np.uint8)
    mats = np.zeros(4, phx.MATERIAL)
    mats["albedo"] = (0.5, 0.25, 0.125, 0.75)
    mats["f2"] = (0.3, 0.7, 0, 0)
    mats["itex"][0] = (-1, -1, -1, -1)   # no textures
    mats["itex"][1] = (5, 9, 2, 0)       # indices beyond the list
    mats["itex"][2] = (0, -1, -1, 0)     # albedo from a 1-byte texture: constant stays
    mats["itex"][3] = (1, 1, 0, 0)       # 2-byte: albedo constant, roughness byte 0;
    #                                      metallic from the 1-byte texture
    texs = [gray, two]
    for m, want in [(0, (0.5, 0.25, 0.125, 0.75, 0.3, 0.7)),
                    (1, (0.5, 0.25, 0.125, 0.75, 0.3, 0.7)),
                    (2, (0.5, 0.25, 0.125, 0.75, 0.3, 0.7)),
                    (3, (0.5, 0.25, 0.125, 0.75, np.float32(90) / np.float32(255),
                         np.float32(77) / np.float32(255)))]:
        d = datas.copy()
        d["mat"] = m
        rec = _records(d, mats, [(0.25, 0.25)], texs)
        got = (*rec["albedo"][0], rec["alpha"][0], rec["roughness"][0], rec["metallic"][0])
        assert np.array_equal(np.float32(got), np.float32(want)), m
    # A material index beyond the table reads the default material, textures or not.
    d = datas.copy()
    d["mat"] = 9
    rec = _records(d, mats, [(0.25, 0.25)], texs)
    assert tuple(rec["albedo"][0]) == (1, 1, 1) and rec["alpha"][0] == 1
    assert rec["roughness"][0] == 1 and rec["metallic"][0] == 0


def test_interpret_wraps_uv_outside_the_unit_square():
    # uv0 offset by (-2, 5): the same texels as inside the unit square.
    inside = _one_triangle((0, 0), (1, 0), (0, 1))
    outside = _one_triangle((-2, 5), (-1, 5), (-2, 6))
    img = np.random.default_rng(1).integers(0, 256, (5, 7, 4), dtype=np.uint8)
    mats = np.zeros(1, phx.MATERIAL)
    mats["itex"] = (0, 0, 0, 0)
    uvs = [((x + 0.5) / 7, (y + 0.5) / 5) for y in range(5) for x in range(7)]
    a = _records(inside, mats, uvs, [img])
    b = _records(outside, mats, uvs, [img])
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert len(np.unique(a["albedo"], axis=0)) > 20


def test_texture_arrays_describe_images():
    """The ctypes PhxTexture records the binding builds: packed (h, w, bpp) and (h, w) arrays, and
    (rows, width, bpp) with an explicit pitch."""
    rgba = np.zeros((3, 5, 4), np.uint8)
    gray = np.zeros((2, 7), np.uint8)
    rows = np.zeros((4, 20), np.uint8)
    arr, keep = phx.texture_arrays([rgba, gray, (rows, 6, 2)])
    got = [(t.width, t.height, t.bytesPerPixel, t.pitch) for t in arr[:3]]
    assert got == [(5, 3, 4, 20), (7, 2, 1, 7), (6, 4, 2, 20)]
    assert all(arr[i].data == keep[i].ctypes.data for i in range(3))
    empty, _ = phx.texture_arrays([])
    assert len(empty) == 1  # one zeroed record, never read with a count of 0
