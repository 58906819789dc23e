"""Seeded synthetic tables for the scan / pre-aggregation / star-join
workloads (numpy PCG64, seed 20261202 + workload id; fragment i of a batch
uses seed + 1000*i).  NULLs are independent Bernoulli per cell."""
import numpy as np

from .kds import ArrowVarlena, Column

SEED_BASE = 20261202
ALNUM = np.frombuffer(b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789",
                      dtype=np.uint8)


def rng(workload, fragment=0):
    return np.random.Generator(np.random.PCG64(SEED_BASE + workload + 1000 * fragment))


def random_text(r, n, lo=8, hi=24, nulls=None):
    """n strings of uniform length lo..hi over [A-Za-z0-9], Arrow-encoded;
    NULL rows get zero length."""
    lens = r.integers(lo, hi + 1, size=n, dtype=np.int64)
    if nulls is not None:
        lens[nulls] = 0
    offs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    data = ALNUM[r.integers(0, len(ALNUM), size=int(offs[-1]))].tobytes()
    return ArrowVarlena(offs, data)


def scan_table(n, seed_workload=0, null_frac=0.05, fragment=0):
    """a int4 U[0,1e6), b float8 N(0,1), c int8 U[-1e12,1e12], d text 8-24
    chars; null_frac NULLs per column."""
    r = rng(seed_workload, fragment)
    a = r.integers(0, 1_000_000, size=n, dtype=np.int32)
    b = r.standard_normal(n)
    c = r.integers(-10**12, 10**12, size=n, dtype=np.int64, endpoint=True)
    nulls = [r.random(n) < null_frac for _ in range(4)]
    d = random_text(r, n, nulls=nulls[3])
    return [Column("a", "int4", a, nulls[0]), Column("b", "float8", b, nulls[1]),
            Column("c", "int8", c, nulls[2]), Column("d", "text", d, nulls[3])]


def zipf_keys(r, n, nkeys=10_000, s=1.1):
    """Bounded Zipf: inverse-CDF sampling from p_k ~ k^-s, k = 1..nkeys."""
    p = np.arange(1, nkeys + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    u = r.random(n)
    return (np.searchsorted(cdf, u, side="right") + 1).astype(np.int32)


def preagg_table(n, seed_workload=1, fragment=0):
    """key int4 Zipf(1.1) over 10,000 values, v int8 U[0,1e6), w float8
    U[0,1000), x int4 uniform over the full int4 range; no NULLs."""
    r = rng(seed_workload, fragment)
    key = zipf_keys(r, n)
    v = r.integers(0, 1_000_000, size=n, dtype=np.int64)
    w = r.random(n) * 1000.0
    x = r.integers(-2**31, 2**31 - 1, size=n, dtype=np.int32, endpoint=True)
    return [Column("key", "int4", key), Column("v", "int8", v),
            Column("w", "float8", w), Column("x", "int4", x)]


def star_tables(n_fact, dim_sizes=(2_000_000, 200_000, 2_556), seed_workload=2):
    """Fact table with int4 foreign keys uniform over each dimension's PK
    range 1..size, quantity int8 U[1,50], price int8 U[1e3,1e7]; dimensions
    with int4 pk 1..size, int2 category U[0,24] and a 16-char text."""
    r = rng(seed_workload)
    fact = [Column("fk%d" % (i + 1), "int4",
                   r.integers(1, sz + 1, size=n_fact, dtype=np.int32))
            for i, sz in enumerate(dim_sizes)]
    fact.append(Column("quantity", "int8", r.integers(1, 51, size=n_fact, dtype=np.int64)))
    fact.append(Column("price", "int8", r.integers(1000, 10_000_001, size=n_fact, dtype=np.int64)))
    dims = []
    for i, sz in enumerate(dim_sizes):
        pk = np.arange(1, sz + 1, dtype=np.int32)
        cat = r.integers(0, 25, size=sz, dtype=np.int16)
        txt = random_text(r, sz, 16, 16)
        dims.append([Column("pk", "int4", pk), Column("category", "int2", cat),
                     Column("txt", "text", txt)])
    return fact, dims


def heap_table(n, seed_workload=5, null_frac=0.1, fragment=0):
    """Mixed-width columns whose heap layout exercises every alignment rule:
    e int2, a int4, f bool, b float8, g text (0-300 bytes: short and 4-byte
    varlena headers), c int8, h float4, d text 8-24 chars; null_frac NULLs
    per column."""
    r = rng(seed_workload, fragment)
    nulls = [r.random(n) < null_frac for _ in range(8)]
    e = r.integers(-2**15, 2**15, size=n, dtype=np.int16)
    a = r.integers(0, 1_000_000, size=n, dtype=np.int32)
    f = r.random(n) < 0.5
    b = r.standard_normal(n)
    glen = np.where(r.random(n) < 0.3, r.integers(127, 301, size=n), r.integers(0, 40, size=n))
    glen[nulls[4]] = 0
    goffs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(glen, out=goffs[1:])
    g = ArrowVarlena(goffs, ALNUM[r.integers(0, len(ALNUM), size=int(goffs[-1]))].tobytes())
    c = r.integers(-10**12, 10**12, size=n, dtype=np.int64, endpoint=True)
    h = r.standard_normal(n).astype(np.float32)
    d = random_text(r, n, nulls=nulls[7])
    return [Column("e", "int2", e, nulls[0]), Column("a", "int4", a, nulls[1]),
            Column("f", "bool", f, nulls[2]), Column("b", "float8", b, nulls[3]),
            Column("g", "text", g, nulls[4]), Column("c", "int8", c, nulls[5]),
            Column("h", "float4", h, nulls[6]), Column("d", "text", d, nulls[7])]


def take_rows(cols, keep):
    """The columns restricted to the rows where keep is True."""
    keep = np.asarray(keep, dtype=bool)
    idx = np.nonzero(keep)[0]
    out = []
    for c in cols:
        nulls = None if c.nulls is None else c.nulls[keep]
        if c.typname in ("text", "bytea", "bpchar"):
            vals = [bytes(c.values[i]) for i in idx]
        else:
            vals = c.values[keep]
        out.append(Column(c.name, c.typname, vals, nulls))
    return out
