"""Partial-result files and their merge, on the CPU (no device needed).

* the reference's own round-trip test (src/test_api.cpp:145-240), replayed
  through write_partial / read_partial / read_partial_header /
  read_partial_one_stripe;
* format compatibility both ways: files this library writes are parsed here
  and every LZ4 block is decoded by the system liblz4 (the codec the
  reference links) and by an independent decoder written from the LZ4 block
  specification; files assembled here with liblz4's LZ4_compress_default (as
  src/api.cpp:1972-2055 writes them) are read back by this library;
* the merge against the reference's stripe-layout goldens
  (src/test_su.cpp:1068-1316: even, odd and odd(2) sample counts, with
  duplicated pairs resolved as the reference does);
* the error codes of validate_partial / merge (src/api.cpp:2322-2394) and of
  the readers.
"""
import ctypes as C
import ctypes.util
import json
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN

MAGIC = 0x088ABA02


# ---------------------------------------------------------------- helpers

def lz4_decode_spec(block: bytes, out_len: int) -> bytes:
    """LZ4 block decoder written from the format specification."""
    out = bytearray()
    i = 0
    while True:
        token = block[i]
        i += 1
        nlit = token >> 4
        if nlit == 15:
            while True:
                b = block[i]
                i += 1
                nlit += b
                if b != 255:
                    break
        out += block[i:i + nlit]
        i += nlit
        if i == len(block):
            break
        off = block[i] | (block[i + 1] << 8)
        i += 2
        mlen = token & 15
        if mlen == 15:
            while True:
                b = block[i]
                i += 1
                mlen += b
                if b != 255:
                    break
        mlen += 4
        assert 0 < off <= len(out)
        for _ in range(mlen):
            out.append(out[-off])
    assert len(out) == out_len
    return bytes(out)


def _liblz4():
    for name in ("liblz4.so.1", ctypes.util.find_library("lz4")):
        if not name:
            continue
        try:
            L = C.CDLL(name)
        except OSError:
            continue
        L.LZ4_compress_default.restype = C.c_int
        L.LZ4_decompress_safe.restype = C.c_int
        L.LZ4_compressBound.restype = C.c_int
        return L
    return None


LZ4 = _liblz4()
need_lz4 = pytest.mark.skipif(LZ4 is None, reason="system liblz4 not present")


def lz4_lib_compress(raw: bytes) -> bytes:
    cap = LZ4.LZ4_compressBound(len(raw))
    dst = C.create_string_buffer(cap)
    n = LZ4.LZ4_compress_default(raw, dst, len(raw), cap)
    assert n > 0
    return dst.raw[:n]


def lz4_lib_decompress(block: bytes, out_len: int) -> bytes:
    dst = C.create_string_buffer(out_len)
    n = LZ4.LZ4_decompress_safe(block, dst, len(block), out_len)
    assert n == out_len
    return dst.raw


def parse_partial(path):
    """(header words, compressed id block, [compressed stripe blocks], footer)."""
    data = open(path, "rb").read()
    h = struct.unpack_from("<8I", data, 0)
    pos = 32
    ids = data[pos:pos + h[7]]
    pos += h[7]
    blocks = []
    for _ in range(h[2]):
        (size,) = struct.unpack_from("<I", data, pos)
        pos += 4
        blocks.append(data[pos:pos + size])
        pos += size
    (footer,) = struct.unpack_from("<I", data, pos)
    assert pos + 4 == len(data)
    return h, ids, blocks, footer


def make_partial(ssu, ids, stripes, start, total, upper=True):
    """A partial_mat_t owned by Python (kept alive by the returned tuple)."""
    arr = [np.ascontiguousarray(s, dtype=np.float64) for s in stripes]
    ptrs = (C.POINTER(C.c_double) * max(1, len(arr)))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arr])
    enc = [i.encode() for i in ids]
    idp = (C.c_char_p * len(enc))(*enc)
    pm = ssu.PartialMat(len(ids), idp, ptrs, start, start + len(arr), total, upper)
    return pm, (arr, ptrs, enc, idp)


def write_reference_style(path, ids, stripes, start, total):
    """A partial file assembled exactly as src/api.cpp:1972-2055 writes one, using liblz4."""
    raw_ids = b"".join(i.encode() + b"\0" for i in ids)
    cids = lz4_lib_compress(raw_ids)
    with open(path, "wb") as f:
        f.write(struct.pack("<8I", MAGIC, len(ids), len(stripes), start, total, 1, len(raw_ids), len(cids)))
        f.write(cids)
        for s in stripes:
            c = lz4_lib_compress(np.asarray(s, np.float64).tobytes())
            f.write(struct.pack("<I", len(c)))
            f.write(c)
        f.write(struct.pack("<I", MAGIC))


def read_dyn(ssu, path):
    p = C.POINTER(ssu.PartialDynMat)()
    st = ssu.lib().read_partial_header(str(path).encode(), C.byref(p))
    return st, p


# ------------------------------------------------------------------ tests

def _test_pm_stripes():
    # src/test_api.cpp make_test_pm(0): 6 samples, 3 stripes, value i*6+j+1
    return ["A", "B", "Cx", "D", "E", "F"], [[i * 6 + j + 1 for j in range(6)] for i in range(3)]


def test_reference_round_trip(ssu, tmp_path):
    """src/test_api.cpp:145-240, step for step."""
    L = ssu.lib()
    ids, stripes = _test_pm_stripes()
    pm, keep = make_partial(ssu, ids, stripes, 0, 3)
    path = str(tmp_path / "ssu_io.dat").encode()
    assert L.write_partial(path, C.byref(pm)) == ssu.WRITE_OKAY

    obs = C.POINTER(ssu.PartialMat)()
    assert L.read_partial(path, C.byref(obs)) == ssu.READ_OKAY
    o = obs.contents
    assert (o.n_samples, o.stripe_start, o.stripe_stop, o.stripe_total) == (6, 0, 3, 3)
    assert [o.sample_ids[i].decode() for i in range(6)] == ids
    for i in range(3):
        assert [o.stripes[i][j] for j in range(6)] == stripes[i]
    L.destroy_partial_mat(C.byref(obs))
    assert not obs

    st, dyn = read_dyn(ssu, tmp_path / "ssu_io.dat")
    assert st == ssu.READ_OKAY
    d = dyn.contents
    assert (d.n_samples, d.stripe_start, d.stripe_stop, d.stripe_total) == (6, 0, 3, 3)
    assert [d.sample_ids[i].decode() for i in range(6)] == ids
    assert not any(bool(d.stripes[i]) for i in range(3))
    for order, loaded in ((1, [False, True, False]), (0, [True, True, False]), (2, [True, True, True])):
        assert L.read_partial_one_stripe(dyn, order) == ssu.READ_OKAY
        assert [bool(d.stripes[i]) for i in range(3)] == loaded
        assert [d.stripes[order][j] for j in range(6)] == stripes[order]
    L.destroy_partial_dyn_mat(C.byref(dyn))
    assert not dyn


@pytest.mark.parametrize("kind", ["repetitive", "random", "constant", "tiny"])
def test_written_blocks_decode_with_liblz4_and_spec(ssu, tmp_path, kind):
    rng = np.random.default_rng(3)
    n = {"repetitive": 4000, "random": 777, "constant": 5000, "tiny": 2}[kind]
    if kind == "repetitive":
        base = rng.random(37)
        stripes = [np.tile(base, n // 37 + 1)[:n] for _ in range(3)]
    elif kind == "random":
        stripes = [rng.random(n) for _ in range(3)]
    elif kind == "constant":
        stripes = [np.full(n, 0.25) for _ in range(3)]
    else:
        stripes = [np.array([0.5, 0.75])]
    total = (n + 1) // 2
    ids = [f"sample.{i}" for i in range(n)]
    pm, keep = make_partial(ssu, ids, stripes, total - len(stripes), total)
    path = tmp_path / "w.partial"
    assert ssu.lib().write_partial(str(path).encode(), C.byref(pm)) == ssu.WRITE_OKAY
    h, cids, blocks, footer = parse_partial(path)
    assert h[0] == MAGIC and footer == MAGIC
    assert tuple(h[1:6]) == (n, len(stripes), total - len(stripes), total, 1)
    raw_ids = b"".join(i.encode() + b"\0" for i in ids)
    assert h[6] == len(raw_ids)
    assert lz4_decode_spec(cids, len(raw_ids)) == raw_ids
    for b, s in zip(blocks, stripes):
        assert lz4_decode_spec(b, 8 * n) == s.tobytes()
        if LZ4 is not None:
            assert lz4_lib_decompress(b, 8 * n) == s.tobytes()
    if kind in ("repetitive", "constant"):
        assert sum(len(b) for b in blocks) < 0.2 * 8 * n * len(stripes)   # it does compress


@need_lz4
def test_reads_files_written_like_the_reference(ssu, tmp_path):
    rng = np.random.default_rng(11)
    n = 301
    ids = [f"S{i:04d}" for i in range(n)]
    stripes = [np.round(rng.random(n), 3) for _ in range(4)]
    path = tmp_path / "ref_style.partial"
    write_reference_style(path, ids, stripes, 7, 151)
    obs = C.POINTER(ssu.PartialMat)()
    assert ssu.lib().read_partial(str(path).encode(), C.byref(obs)) == ssu.READ_OKAY
    o = obs.contents
    assert (o.n_samples, o.stripe_start, o.stripe_stop, o.stripe_total, o.is_upper_triangle) == (n, 7, 11, 151, True)
    assert [o.sample_ids[i].decode() for i in range(n)] == ids
    for i in range(4):
        assert np.array_equal(np.ctypeslib.as_array(o.stripes[i], shape=(n,)), stripes[i])
    ssu.lib().destroy_partial_mat(C.byref(obs))


def _merge(ssu, tmp_path, ids, stripes, splits, fp32=False, mmap_dir=None, preload=()):
    """preload: (partial, stripe) pairs read with read_partial_one_stripe before the merge."""
    L = ssu.lib()
    total = len(stripes)
    mats = []
    for a, b in zip(splits[:-1], splits[1:]):
        pm, keep = make_partial(ssu, ids, stripes[a:b], a, total)
        path = tmp_path / f"p{a}.partial"
        assert L.write_partial(str(path).encode(), C.byref(pm)) == ssu.WRITE_OKAY
        st, dyn = read_dyn(ssu, path)
        assert st == ssu.READ_OKAY
        mats.append(dyn)
    arr = (C.POINTER(ssu.PartialDynMat) * len(mats))(*mats)
    assert L.validate_partial(arr, len(mats)) == ssu.MERGE_OKAY
    for p, q in preload:
        assert L.read_partial_one_stripe(mats[p], q) == ssu.READ_OKAY
    n = len(ids)
    md = None if mmap_dir is None else str(mmap_dir).encode()
    if fp32:
        res = C.POINTER(ssu.MatFullFP32)()
        st = L.merge_partial_to_mmap_matrix_fp32(arr, len(mats), md, C.byref(res))
    else:
        res = C.POINTER(ssu.MatFullFP64)()
        st = L.merge_partial_to_mmap_matrix(arr, len(mats), md, C.byref(res))
    assert st == ssu.MERGE_OKAY
    m = np.ctypeslib.as_array(res.contents.matrix, shape=(n * n,)).copy().reshape(n, n)
    got_ids = [res.contents.sample_ids[i].decode() for i in range(n)]
    (L.destroy_mat_full_fp32 if fp32 else L.destroy_mat_full_fp64)(C.byref(res))
    for d in mats:
        L.destroy_partial_dyn_mat(C.byref(d))
    assert got_ids == ids
    return m


@pytest.mark.parametrize("case", ["stripes_to_matrix_even", "stripes_to_matrix_odd", "stripes_to_matrix_odd2"])
@pytest.mark.parametrize("fp32", [False, True])
def test_merge_matches_reference_layout_goldens(ssu, tmp_path, case, fp32):
    g = json.load(open(os.path.join(GOLDEN, "stripe_layouts.json")))[case]
    n = g["n_samples"]
    ids = [f"s{i}" for i in range(n)]
    m = _merge(ssu, tmp_path, ids, g["stripes"], [0, 2, g["n_stripes"]], fp32=fp32,
               mmap_dir=tmp_path if fp32 else None)
    assert np.array_equal(m.ravel(), np.asarray(g["expected"], np.float32 if fp32 else np.float64))


@pytest.mark.parametrize("case", ["stripes_to_condensed_form_even", "stripes_to_condensed_form_odd",
                                  "stripes_to_condensed_form_odd2"])
def test_oracle_condensed_form_matches_reference_goldens(O, case):
    g = json.load(open(os.path.join(GOLDEN, "stripe_layouts.json")))[case]
    n, ns = g["n_samples"], g["n_stripes"]
    got = O.stripes_to_condensed(np.asarray(g["stripes"], np.float64)[:ns], n, 0)
    assert np.array_equal(got, np.asarray(g["expected"]))


def test_merge_of_random_matrix_stripes(ssu, O, tmp_path):
    """Stripes cut from a symmetric matrix merge back into it, for several splits and both parities."""
    rng = np.random.default_rng(5)
    for n in (7, 8, 64, 65):
        a = rng.random((n, n))
        full = a + a.T
        np.fill_diagonal(full, 0)
        total = (n + 1) // 2
        stripes = [[full[k, (k + s + 1) % n] for k in range(n)] for s in range(total)]
        ids = [f"x{i}" for i in range(n)]
        for splits in ([0, total], [0, 1, total], [0, total // 2, total - 1, total]):
            splits = sorted(set(splits))
            assert np.array_equal(_merge(ssu, tmp_path, ids, stripes, splits), full)


@pytest.mark.parametrize("n", [3000, 3001])
@pytest.mark.parametrize("fp32", [False, True])
def test_merge_on_several_threads(ssu, tmp_path, n, fp32):
    """1,500 stripes in 24 blocks of 64, placed by several threads, with stripes a caller
    loaded beforehand (first, middle and last of a partial), into memory and into an mmap'd file."""
    rng = np.random.default_rng(n)
    a = rng.random((n, n))
    full = a + a.T
    np.fill_diagonal(full, 0)
    total = (n + 1) // 2
    k = np.arange(n)
    stripes = [full[k, (k + s + 1) % n] for s in range(total)]
    ids = [f"x{i}" for i in range(n)]
    splits = [0, 130, 700, 701, total]
    pre = [(0, 0), (1, 63), (1, 64), (3, total - 702)]
    exp = full.astype(np.float32) if fp32 else full
    assert np.array_equal(_merge(ssu, tmp_path, ids, stripes, splits, fp32, preload=pre), exp)
    mm = tmp_path / "mm"
    mm.mkdir()
    assert np.array_equal(_merge(ssu, tmp_path, ids, stripes, splits, fp32, mmap_dir=mm), exp)


def test_merge_of_a_file_cut_after_its_header(ssu, tmp_path):
    """A partial cut short after read_partial_header: the merge reports an incomplete
    stripe set and hands back no result."""
    L = ssu.lib()
    n = 1000
    rng = np.random.default_rng(11)
    total = (n + 1) // 2
    stripes = [rng.random(n) for _ in range(total)]
    ids = [f"x{i}" for i in range(n)]
    mats = []
    for a, b in ((0, 200), (200, total)):
        pm, keep = make_partial(ssu, ids, stripes[a:b], a, total)
        path = tmp_path / f"p{a}.partial"
        assert L.write_partial(str(path).encode(), C.byref(pm)) == ssu.WRITE_OKAY
        st, dyn = read_dyn(ssu, path)
        assert st == ssu.READ_OKAY
        mats.append(dyn)
    last = tmp_path / "p200.partial"
    last.write_bytes(last.read_bytes()[: last.stat().st_size * 3 // 5])
    arr = (C.POINTER(ssu.PartialDynMat) * 2)(*mats)
    res = C.POINTER(ssu.MatFullFP64)()
    assert L.merge_partial_to_mmap_matrix(arr, 2, None, C.byref(res)) == ssu.INCOMPLETE_STRIPE_SET
    assert not res
    for d in mats:
        L.destroy_partial_dyn_mat(C.byref(d))


def test_validate_and_merge_error_codes(ssu, tmp_path):
    """src/api.cpp:2322-2394: the codes of check_partial, in its order."""
    L = ssu.lib()
    ids, _ = _test_pm_stripes()

    def dyn(name, start, stop, total, sids=ids):
        pm, keep = make_partial(ssu, sids, [[1.0] * len(sids)] * (stop - start), start, total)
        path = tmp_path / name
        assert L.write_partial(str(path).encode(), C.byref(pm)) == ssu.WRITE_OKAY
        st, d = read_dyn(ssu, path)
        assert st == ssu.READ_OKAY
        return d

    def check(mats):
        arr = (C.POINTER(ssu.PartialDynMat) * len(mats))(*mats)
        v = L.validate_partial(arr, len(mats))
        res = C.POINTER(ssu.MatFullFP64)()
        assert L.merge_partial_to_mmap_matrix(arr, len(mats), None, C.byref(res)) == v
        if v == ssu.MERGE_OKAY:
            L.destroy_mat_full_fp64(C.byref(res))
        return v

    a, b, c = dyn("a", 0, 2, 5), dyn("b", 2, 4, 5), dyn("c", 4, 5, 5)
    assert check([a, b, c]) == ssu.MERGE_OKAY
    assert check([c, a, b]) == ssu.MERGE_OKAY                     # any order
    assert check([a, c]) == ssu.INCOMPLETE_STRIPE_SET
    assert check([a, dyn("b2", 1, 4, 5), c]) == ssu.STRIPES_OVERLAP
    other = list(ids)
    other[2] = "X"
    assert check([a, dyn("b3", 2, 4, 5, other), c]) == ssu.SAMPLE_ID_CONSISTENCY
    assert check([a, b, dyn("c2", 3, 4, 4)]) == ssu.PARTIALS_MISMATCH                  # stripe_total
    assert check([a, b, dyn("c3", 4, 5, 5, ids + ["G"])]) == ssu.PARTIALS_MISMATCH   # sample count
    for d in (a, b, c):
        L.destroy_partial_dyn_mat(C.byref(d))


def test_reader_error_codes(ssu, tmp_path):
    L = ssu.lib()
    obs = C.POINTER(ssu.PartialMat)()
    assert L.read_partial(str(tmp_path / "missing").encode(), C.byref(obs)) == ssu.OPEN_ERROR
    ids, stripes = _test_pm_stripes()
    pm, keep = make_partial(ssu, ids, stripes, 0, 3)
    good = tmp_path / "good.partial"
    assert L.write_partial(str(good).encode(), C.byref(pm)) == ssu.WRITE_OKAY
    data = good.read_bytes()
    bad_magic = tmp_path / "bad_magic"
    bad_magic.write_bytes(b"\0\0\0\0" + data[4:])
    assert L.read_partial(str(bad_magic).encode(), C.byref(obs)) == ssu.MAGIC_INCOMPATIBLE
    bad_hdr = tmp_path / "bad_header"      # stripe_total >= n_samples
    bad_hdr.write_bytes(data[:16] + struct.pack("<I", 6) + data[20:])
    assert L.read_partial(str(bad_hdr).encode(), C.byref(obs)) == ssu.BAD_HEADER
    lower = tmp_path / "lower"             # is_upper_triangle = 0
    lower.write_bytes(data[:20] + struct.pack("<I", 0) + data[24:])
    assert L.read_partial(str(lower).encode(), C.byref(obs)) == ssu.BAD_HEADER
    cut = tmp_path / "truncated"
    cut.write_bytes(data[:-10])
    assert L.read_partial(str(cut).encode(), C.byref(obs)) == ssu.MAGIC_INCOMPATIBLE
    nofoot = tmp_path / "no_footer"
    nofoot.write_bytes(data[:-4] + b"\1\2\3\4")
    assert L.read_partial(str(nofoot).encode(), C.byref(obs)) == ssu.MAGIC_INCOMPATIBLE
    assert not obs
    assert L.write_partial(str(tmp_path / "no_dir" / "x").encode(), C.byref(pm)) == ssu.WRITE_ERROR
