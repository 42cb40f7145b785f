"""Signature-file codec (include/xjpcg_sig.h) on the CPU.

Pins, in order of strength:
  * byte identity with the reference's own partitioner: against the committed
    fixture (tests/golden/signature_cases.npz, written by the reference's
    SpmPar + Signature::store_*), and, where oracle/_ref is built, against the
    reference run live on random matrices, including its value-only update;
  * decoding equals the oracle's restatement of the reference decoder
    (oracle.sig_decode) entry for entry, in stream order;
  * decode(encode(A)) returns A's non-zero entries (a property at sizes the
    reference's single-threaded encoder would take long on);
  * malformed files, bad arguments and the reference's partitioner errors
    return the documented statuses.
"""
import os

import numpy as np
import pytest

from hpc_amd._lib import XjpcgError
from hpc_amd.signature import Signature
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "signature_cases.npz")
INVALID_VALUE, INTERNAL_ERROR = 3, 5


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN) as z:
        return {k: z[k] for k in z.files}


def case_files(g, name):
    return {fn: g[f"{name}/{fn}"].tobytes() for fn in oracle.SIG_FILES}


def write_files(directory, files):
    for fn, b in files.items():
        with open(os.path.join(directory, fn), "wb") as f:
            f.write(b)


def encode_case(g, name):
    if name == "cscsym":
        s = Signature.from_cscsym(int(g[f"{name}/n"]), g[f"{name}/row_idx"], g[f"{name}/col_ptr"], g[f"{name}/val"],
                                  int(g[f"{name}/base"]))
        s.update(g[f"{name}/val2"])
        return s
    return Signature.from_coo(int(g[f"{name}/m"]), int(g[f"{name}/n"]), g[f"{name}/row"], g[f"{name}/col"],
                              g[f"{name}/val"], int(g[f"{name}/base"]))


def expected_entries(g, name):
    """0-based non-zero entries of a case's matrix, lexsorted (row, col, val)."""
    if name == "cscsym":
        n = int(g["cscsym/n"])
        cp, ri, v = g["cscsym/col_ptr"], g["cscsym/row_idx"], g["cscsym/val2"]
        cols = np.repeat(np.arange(n), np.diff(cp))
        off = ri != cols
        r = np.concatenate([ri, cols[off]])
        c = np.concatenate([cols, ri[off]])
        v = np.concatenate([v, v[off]])
    else:
        b = int(g[f"{name}/base"])
        r = g[f"{name}/row"].astype(np.int64) - b
        c = g[f"{name}/col"].astype(np.int64) - b
        v = g[f"{name}/val"]
    keep = v != 0
    return lexsorted(r[keep], c[keep], v[keep])


def lexsorted(r, c, v):
    r = np.asarray(r, np.int64)
    c = np.asarray(c, np.int64)
    v = np.asarray(v, np.float64)
    o = np.lexsort((v, c, r))
    return r[o], c[o], v[o]


def assert_same_files(a, b):
    bad = [fn for fn in oracle.SIG_FILES if a[fn] != b[fn]]
    assert not bad, f"files differ: {bad}"


CASES = ["stencil", "rect", "longrow", "cscsym"]


@pytest.mark.parametrize("name", CASES)
def test_golden_bytes(golden, name, tmp_path):
    s = encode_case(golden, name)
    s.store(str(tmp_path))
    assert_same_files(oracle.read_sig_files(str(tmp_path)), case_files(golden, name))
    assert s.buffer("rbParam") == case_files(golden, name)["rbParam.dat"]
    assert s.buffer(5) == case_files(golden, name)["nnzVal_5.dat"]


@pytest.mark.parametrize("name", CASES)
def test_golden_decode(golden, name, tmp_path):
    files = case_files(golden, name)
    write_files(str(tmp_path), files)
    s = Signature.load(str(tmp_path))
    r, c, v = s.to_coo()
    ro, co, vo = oracle.sig_decode(files)
    np.testing.assert_array_equal(r, ro)
    np.testing.assert_array_equal(c, co)
    np.testing.assert_array_equal(v, vo)
    er, ec, ev = expected_entries(golden, name)
    gr, gc, gv = lexsorted(r, c, v)
    np.testing.assert_array_equal(gr, er)
    np.testing.assert_array_equal(gc, ec)
    np.testing.assert_array_equal(gv, ev)
    info = s.info()
    hdr = np.frombuffer(files["info.dat"], np.uint32)
    assert [info[k] for k in ("m", "n", "nnz", "m_pad", "n_pad", "nnz_pad")] == hdr.tolist()
    assert info["entries"] == er.size


def test_golden_structure(golden):
    """Counts the format fixes, read straight from the reference's bytes."""
    files = case_files(golden, "rect")
    m, n, nnz, m_pad, n_pad, nnz_pad = np.frombuffer(files["info.dat"], np.uint32).tolist()
    assert (m, n, nnz, m_pad) == (9000, 13000, 600, 9000)
    assert n_pad == -(-n // 4) * 4
    rb = np.frombuffer(files["rbParam.dat"], np.uint32).reshape(-1, 8)
    par = np.frombuffer(files["parParam.dat"], np.uint32).reshape(-1, 8)
    n_rb, n_par = rb[0, 1], par[0, 0]
    assert rb.shape[0] == 1 + 8 * n_rb and par.shape[0] == 1 + 8 * n_par
    assert rb[0, 0] == m
    blocks = rb[1:].reshape(n_rb, 8, 8)
    assert blocks[:, 0, 3].sum() == n_par                        # partitions per block
    assert np.all(np.diff(blocks[:, 0, 0]) >= 4096)              # blocks start >= 4096 rows apart
    assert np.all(blocks[:, 1, 0] <= 4096)
    parts = par[1:].reshape(n_par, 8, 8)
    ch_nnz = parts[:, 2:4, :].reshape(n_par, 16)
    assert np.all(ch_nnz % 32 == 0)                              # channels hold whole padded rows
    assert np.array_equal(ch_nnz.sum(axis=1), parts[:, 4, 3])    # channel counts add to the partition's
    assert parts[:, 4, 3].sum() == nnz_pad
    assert np.all(parts[:, 4, 1] <= 1024)                        # <= 4096 columns = 1024 blocks
    for c in range(16):
        st = np.frombuffer(files[f"nnzVal_{c}.dat"], np.uint32)
        tot, rw, cw, vw = st[:4]
        assert tot == rw + cw + vw and st.size * 4 == 32 * (1 + tot)
        cnt = ch_nnz[:, c].astype(np.int64)
        assert vw == (cnt // 4).sum()
        assert cw == (-(-cnt // 64)).sum() and rw == (-(-cnt // 512)).sum()


def random_coo(rng, m, n, k, base=0, dup=0, zeros=0):
    r = rng.integers(0, m, k).astype(np.uint32)
    c = rng.integers(0, n, k).astype(np.uint32)
    if dup:
        r[:dup] = r[k - dup:]
        c[:dup] = c[k - dup:]
    v = rng.standard_normal(k)
    if zeros:
        v[::zeros] = 0.0
    return r + base, c + base, v


REF = oracle.ref_lib() is not None
need_ref = pytest.mark.skipif(not REF, reason="oracle/_ref (the reference's partitioner) is not built")


@need_ref
@pytest.mark.parametrize("seed", range(6))
def test_reference_bytes_random(seed, tmp_path):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(1, 20000))
    n = int(rng.integers(1, 20000))
    k = int(rng.integers(1, 4000))
    base = seed % 2
    r, c, v = random_coo(rng, m, n, k, base, dup=min(k // 4, 50), zeros=7)
    a, b = tmp_path / "mine", tmp_path / "ref"
    a.mkdir()
    b.mkdir()
    Signature.from_coo(m, n, r, c, v, base).store(str(a))
    oracle.ref_sig_coo(str(b), m, n, r, c, v, base=base)
    assert_same_files(oracle.read_sig_files(str(a)), oracle.read_sig_files(str(b)))


@need_ref
@pytest.mark.parametrize("shape", [(4097, 4097, 12000), (100, 100000, 8000), (70000, 50, 6000), (1, 1, 1)])
def test_reference_bytes_shapes(shape, tmp_path):
    """Block and partition edges: 4097 rows (a one-row second block), one very wide row range,
    a tall thin matrix, a single entry; each followed by the value-only update."""
    m, n, k = shape
    rng = np.random.default_rng(m + n + k)
    r, c, v = random_coo(rng, m, n, k, dup=min(k // 3, 200))
    if m == 4097:
        r[-1] = 4096
    v2 = rng.standard_normal(k)
    a, b = tmp_path / "mine", tmp_path / "ref"
    a.mkdir()
    b.mkdir()
    s = Signature.from_coo(m, n, r, c, v)
    s.update(v2)
    s.store(str(a))
    oracle.ref_sig_coo(str(b), m, n, r, c, v, val2=v2)
    assert_same_files(oracle.read_sig_files(str(a)), oracle.read_sig_files(str(b)))


def test_round_trip_large():
    """decode(encode(A)) = A's non-zero entries on a 60000 x 60000 matrix with 400k entries
    (duplicates and zero values included) and on a 2D 5-point matrix on a 300 x 300 grid."""
    rng = np.random.default_rng(7)
    r, c, v = random_coo(rng, 60000, 60000, 400000, dup=5000, zeros=11)
    s = Signature.from_coo(60000, 60000, r, c, v)
    keep = v != 0
    got = lexsorted(*s.to_coo())
    exp = lexsorted(r[keep], c[keep], v[keep])
    for a, b in zip(got, exp):
        np.testing.assert_array_equal(a, b)
    info = s.info()
    assert info["nnz_pad"] % 32 == 0 and info["entries"] == keep.sum()

    g = 300
    i, j = np.divmod(np.arange(g * g), g)
    rows, cols = [], []
    for di, dj in ((0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)):
        ok = (i + di >= 0) & (i + di < g) & (j + dj >= 0) & (j + dj < g)
        rows.append((i * g + j)[ok])
        cols.append(((i + di) * g + j + dj)[ok])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = rng.uniform(0.5, 1.5, r.size)
    s = Signature.from_coo(g * g, g * g, r, c, v)
    got = lexsorted(*s.to_coo())
    exp = lexsorted(r, c, v)
    for a, b in zip(got, exp):
        np.testing.assert_array_equal(a, b)
    assert s.info()["row_blocks"] == -(-g * g // 4096)


def test_update_values(golden):
    """Value-only update rewrites only value words; decoding returns the new values."""
    g = golden
    s = encode_case(g, "rect")
    before = {c: s.buffer(c) for c in range(16)}
    k = g["rect/val"].size
    v2 = np.random.default_rng(1).standard_normal(k)
    s.update(v2)
    b = int(g["rect/base"])
    exp = lexsorted(g["rect/row"].astype(np.int64) - b, g["rect/col"].astype(np.int64) - b, v2)
    got = lexsorted(*s.to_coo())
    for x, y in zip(got, exp):
        np.testing.assert_array_equal(x, y)
    # index words are untouched: bytes outside value words equal
    for c in range(16):
        a = np.frombuffer(before[c], np.uint8).reshape(-1, 32)
        bb = np.frombuffer(s.buffer(c), np.uint8).reshape(-1, 32)
        assert a.shape == bb.shape
        changed = np.any(a != bb, axis=1)
        assert not changed[0]


def status_of(fn, *a):
    with pytest.raises(XjpcgError) as e:
        fn(*a)
    return e.value.status, e.value.message


def test_errors_arguments(tmp_path):
    r = np.array([0, 1], np.uint32)
    c = np.array([0, 5], np.uint32)
    v = np.array([1.0, 2.0])
    st, msg = status_of(Signature.from_coo, 4, 4, r, c, v)
    assert st == INVALID_VALUE and "out of range at entry 1" in msg
    st, _ = status_of(Signature.from_coo, 0, 4, r, r, v)
    assert st == INVALID_VALUE
    st, msg = status_of(Signature.from_coo, 4, 4, r[:0], r[:0], v[:0])
    assert st == INTERNAL_ERROR
    st, msg = status_of(Signature.from_coo, 4, 4, r, r, v, 2)
    assert st == INTERNAL_ERROR and "unsupported storage mode" in msg
    s = Signature.from_coo(4, 4, r, r, v)
    st, msg = status_of(s.update, np.ones(3))
    assert st == INTERNAL_ERROR and "updateing matrix" in msg
    st, msg = status_of(s.store, str(tmp_path / "missing" / "dir"))
    assert st == INVALID_VALUE and "cannot open" in msg
    s.store(str(tmp_path))
    loaded = Signature.load(str(tmp_path))
    st, msg = status_of(loaded.update, v)
    assert st == INVALID_VALUE and "no value map" in msg
    # CSC-sym whose column pointers do not expand to 2 nnz_half - n entries (reference's loader message)
    st, msg = status_of(Signature.from_cscsym, 3, np.array([0, 1, 2, 2]), np.array([0, 2, 3, 4]), np.ones(4))
    assert st == INTERNAL_ERROR and "loadCscSym" in msg


def test_errors_files(golden, tmp_path):
    files = case_files(golden, "rect")

    def load_with(mutate):
        d = tmp_path / f"case{len(os.listdir(tmp_path))}"
        d.mkdir()
        f = dict(files)
        mutate(f)
        write_files(str(d), f)
        return status_of(Signature.load, str(d))

    st, msg = load_with(lambda f: f.update({"nnzVal_3.dat": f["nnzVal_3.dat"][:-32]}))
    assert st == INVALID_VALUE and "nnzVal_3.dat" in msg
    st, msg = load_with(lambda f: f.update({"nnzVal_4.dat": f["nnzVal_4.dat"] + bytes(32)}))
    assert st == INVALID_VALUE and "longer" in msg
    st, msg = load_with(lambda f: f.update({"info.dat": f["info.dat"][:20]}))
    assert st == INVALID_VALUE and "info.dat" in msg
    st, msg = load_with(lambda f: f.update({"parParam.dat": f["parParam.dat"][:-32]}))
    assert st == INVALID_VALUE and "parParam.dat" in msg

    def bad_header(f):
        b = bytearray(f["nnzVal_0.dat"])
        b[4] ^= 1
        f["nnzVal_0.dat"] = bytes(b)

    st, msg = load_with(bad_header)
    assert st == INVALID_VALUE and "header" in msg

    def small_n(f):
        b = bytearray(f["info.dat"])
        b[4:8] = np.uint32(100).tobytes()
        f["info.dat"] = bytes(b)

    d = tmp_path / "small_n"
    d.mkdir()
    f = dict(files)
    small_n(f)
    write_files(str(d), f)
    s = Signature.load(str(d))
    st, msg = status_of(s.to_coo)
    assert st == INVALID_VALUE and "outside" in msg
    st, msg = status_of(Signature.load, str(tmp_path / "nowhere"))
    assert st == INVALID_VALUE and "info.dat" in msg


def test_to_csr_rejects_rectangular():
    """Checked before any device work, so it runs without a GPU."""
    from hpc_amd._lib import lib
    s = Signature.from_coo(10, 20, np.array([0, 9]), np.array([19, 0]), np.array([1.0, 2.0]))
    with pytest.raises(ValueError):
        s.to_csr()
    st = lib().xjpcg_sig_to_csr(s._handle(), None, None, None, 0, None, None)
    assert st == INVALID_VALUE and "not square" in lib().xjpcg_sig_last_error().decode()
