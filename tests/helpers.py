"""Shared helpers of the test suite: random replay programs and block checks."""
import numpy as np

from halloc_amd.workloads import graph_sizes, sweep_sizes  # noqa: F401  (re-exported for the tests)

BLOCK_SIZES = [16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072]
REC_NULL = 0xFFFFFFFF

def random_ops(rng, nsteps, lanes, sizes, p_free=0.4, large_every=0):
    """Random replay program: each step a lane mallocs, frees a live block, or idles."""
    ops = np.zeros((nsteps, lanes, 2), dtype=np.uint32)
    live = []
    for st in range(nsteps):
        freed_now = set()
        for l in range(lanes):
            u = rng.random()
            if live and u < p_free:
                j = int(rng.integers(len(live)))
                idx = live.pop(j)
                ops[st, l] = (2, idx)
            elif u < 0.9:
                n = int(rng.choice(sizes))
                if large_every and rng.integers(large_every) == 0:
                    n = int(rng.integers(3073, 65536))
                ops[st, l] = (1, n)
                freed_now.add(st * lanes + l)
        live.extend(sorted(freed_now))
    return ops


def live_intervals(ops, recs, table):
    """(start, end) byte offsets within the arena of live small blocks after a replay."""
    nsteps, lanes = ops.shape[:2]
    alive = {}
    for st in range(nsteps):
        for l in range(lanes):
            i = st * lanes + l
            k, a = ops[st, l]
            if k == 1:
                r = recs[st, l]
                if r[0] < 16:
                    alive[i] = (int(r[1]), int(r[2]), int(table[r[0]][11]), int(table[r[0]][8]))
            elif k == 2:
                alive.pop(int(a), None)
    return list(alive.values())


def assert_disjoint(blocks, sb_sz):
    iv = sorted((sb * sb_sz + ch * cb, sb * sb_sz + ch * cb + bs, cb) for sb, ch, cb, bs in blocks)
    for (s0, e0, _), (s1, _, _) in zip(iv, iv[1:]):
        assert e0 <= s1, "overlapping live blocks"
    for s, e, cb in iv:
        assert s % (8 if cb % 3 == 0 else 16) == 0  # 24 B classes are 8-aligned (src/halloc.cu:377-380)




def pressure_program(seed=3, nsteps=6500):
    """Three phases on one wave: fill a small pool with 64/128 B blocks until
    hamalloc returns NULL, free almost everything, then allocate 16/32 B
    blocks, which must reuse slabs first owned by the other classes."""
    rng = np.random.default_rng(seed)
    steps, live = [], []
    for st in range(nsteps):
        row = np.zeros((64, 2), dtype=np.uint32)
        pf, sizes = (0.05, (64, 128)) if st < 2500 else (0.98, (64, 128)) if st < 5500 else (0.3, (16, 32))
        for l in range(64):
            if live and rng.random() < pf:
                row[l] = (2, live.pop(int(rng.integers(len(live)))))
            else:
                row[l] = (1, sizes[int(rng.integers(2))])
        steps.append(row)
        live.extend(st * 64 + l for l in range(64) if row[l, 0] == 1)
    return np.stack(steps)


def crowded_program(seed=5, size=256, fill_steps=100, churn_steps=200):
    """One wave: fill a head with `size` B blocks to about 78% of a 2 MiB slab
    (64 lanes x fill_steps), then churn: every step frees 32 random live blocks
    and allocates 32, so the head keeps its fill while its free blocks scatter
    and a group's probes need many rounds (the crowded-head rule)."""
    rng = np.random.default_rng(seed)
    lanes = 64
    ops = np.zeros((fill_steps + churn_steps, lanes, 2), dtype=np.uint32)
    live = []
    for st in range(fill_steps):
        ops[st, :, 0] = 1
        ops[st, :, 1] = size
        live.extend(range(st * lanes, (st + 1) * lanes))
    for st in range(fill_steps, fill_steps + churn_steps):
        new = []
        for l in range(lanes):
            if l % 2 == 0:
                ops[st, l] = (2, live.pop(int(rng.integers(len(live)))))
            else:
                ops[st, l] = (1, size)
                new.append(st * lanes + l)
        live.extend(new)
    return ops


def fill_program(size, nsteps, lanes=64):
    """Every lane of one wave allocates `size` B in every step."""
    ops = np.zeros((nsteps, lanes, 2), dtype=np.uint32)
    ops[:, :, 0] = 1
    ops[:, :, 1] = size
    return ops


def mixed_sizes_tail(recs, big=128):
    """Continues a fill of 16 B blocks (1 chunk each) whose records are recs
    (steps, lanes, 4): frees every block of the first slab but two, at chunks
    c1 and c2 off the multiples of 8 (c2 at 1..7 mod 16 and not c1 + 8);
    then one lane allocates `big` B (8 chunks, same chunk size), which takes
    the now sparse slab as its head; then lanes 0 and 1 of one wave free the
    two blocks: two runs of one block whose starts are not multiples of the
    slab's block length.  Returns (ops, slab, step of the `big` malloc)."""
    nsteps, lanes = recs.shape[:2]
    sb = recs[:, :, 1].reshape(-1)
    ch = recs[:, :, 2].reshape(-1).astype(np.int64)
    inx = np.nonzero(sb == sb[0])[0]
    k1 = next(int(i) for i in inx if ch[i] % 8)
    k2 = next(int(i) for i in inx if 1 <= ch[i] % 16 <= 7 and ch[i] != ch[k1] + 8 and i != k1)
    frees = [int(i) for i in inx if i not in (k1, k2)]
    nfree = (len(frees) + lanes - 1) // lanes
    ops = np.zeros((nfree + 2, lanes, 2), dtype=np.uint32)
    for j, i in enumerate(frees):
        ops[j // lanes, j % lanes] = (2, i)
    ops[nfree, 0] = (1, big)
    ops[nfree + 1, 0] = (2, k1)
    ops[nfree + 1, 1] = (2, k2)
    return ops, int(sb[0]), nsteps + nfree


def random_batch_ops(rng, nsteps, lanes, sizes, max_count=16, p_free=0.4):
    """Random batched replay program: [kind, arg, count, 0] per (step, lane)."""
    ops = np.zeros((nsteps, lanes, 4), dtype=np.uint32)
    live = []
    for st in range(nsteps):
        new = []
        for l in range(lanes):
            u = rng.random()
            if live and u < p_free:
                ops[st, l] = (2, live.pop(int(rng.integers(len(live)))), 0, 0)
            elif u < 0.9:
                ops[st, l] = (1, int(rng.choice(sizes)), int(rng.integers(1, max_count + 1)), 0)
                new.append(st * lanes + l)
        live.extend(new)
    return ops


def wave_refree_ops(lanes, sizes):
    """Batched replay program, per size: the wave allocates 16 blocks per lane
    and ragged counts; frees the 16-block op keeping its pointers (kind 3);
    frees it again (kind 2) while its slab still holds the ragged op; frees
    the ragged op; then a partial wave allocates and frees.  Every free frees
    a lane's own blocks, so each round is one slab's runs."""
    steps = []
    for size in sizes:
        base = len(steps)
        op = lambda st, l: (base + st) * lanes + l
        prog = [[(1, size, 16, 0)] * lanes,
                [(1, size, 1 + l % 7, 0) for l in range(lanes)],
                [(3, op(0, l), 0, 0) for l in range(lanes)],
                [(2, op(0, l), 0, 0) for l in range(lanes)],
                [(2, op(1, l), 0, 0) for l in range(lanes)],
                [(1, size, 5, 0) if l % 3 == 0 else (0, 0, 0, 0) for l in range(lanes)],
                [(2, op(5, l), 0, 0) if l % 3 == 0 else (0, 0, 0, 0) for l in range(lanes)]]
        steps.extend(prog)
    return np.array(steps, dtype=np.uint32)


def _hash3(a, b, c):
    """The drivers' 32-bit mixing hash (halloc_amd/csrc/harness.hip hash3), on uint32 arrays."""
    u = np.uint32
    a, b, c = (np.atleast_1d(np.asarray(v, dtype=u)) for v in (a, b, c))
    x = (a * u(0x9E3779B1)) ^ ((b + u(0x7F4A7C15)) * u(0x85EBCA77)) ^ ((c + u(0x165667B1)) * u(0xC2B2AE3D))
    x ^= x >> u(16)
    x *= u(0x7FEB352D)
    x ^= x >> u(15)
    x *= u(0x846CA68B)
    x ^= x >> u(16)
    return x


def serve_checksums(tag, produced):
    """Expected per-request checksum of the serving loop: the sum (mod 2^32)
    of the hashes of tokens 0..produced-1."""
    tag = np.asarray(tag, dtype=np.uint32)
    n = np.asarray(produced, dtype=np.int64)
    out = np.zeros(len(n), dtype=np.uint32)
    nz = np.nonzero(n)[0]
    if len(nz) == 0:
        return out
    cnt = n[nz]
    starts = np.concatenate(([0], np.cumsum(cnt)[:-1]))
    t = np.arange(cnt.sum(), dtype=np.int64) - np.repeat(starts, cnt)
    h = _hash3(np.repeat(tag[nz], cnt), t.astype(np.uint32), np.uint32(0x5E7A11C3)).astype(np.uint64)
    out[nz] = (np.add.reduceat(h, starts) & 0xFFFFFFFF).astype(np.uint32)
    return out


# --------------------------------------------------------- large region
LARGE_UNIT = 4096
LARGE_NCLASSES = 40
LARGE_SB = 0xFFFFFFFE


def large_units(memory, halloc_fraction=0.75):
    """4 KiB units of the large region: the pool outside the slab share."""
    return min((memory - int(halloc_fraction * memory)) // LARGE_UNIT, 0xFFFFFFF0)


def large_class_of(nbytes):
    """Class of a large request: the least c with 2^c units of 4 KiB >= nbytes."""
    units = -(-int(nbytes) // LARGE_UNIT)
    return (units - 1).bit_length()


def large_state(dump, lunits):
    """(tickets, unit bits, start words) of the large region: the last words of a dump."""
    lwords = (lunits + 63) // 64
    d = np.asarray(dump)
    lstart = d[len(d) - lunits:]
    lbits = d[len(d) - lunits - lwords:len(d) - lunits]
    tick = d[len(d) - lunits - lwords - LARGE_NCLASSES:len(d) - lunits - lwords]
    return tick, lbits, lstart


def unit_used(lbits, u):
    return bool((int(lbits[u // 64]) >> (u % 64)) & 1)


def live_large_blocks(ops, recs):
    """(first unit, units) of the large blocks live after a replay (kind 2 or 3 frees
    by record: only a free whose record names units takes a block away)."""
    nsteps, lanes = ops.shape[:2]
    alive = {}
    for st in range(nsteps):
        for l in range(lanes):
            i = st * lanes + l
            k, a = ops[st, l]
            r = recs[st, l]
            if k == 1 and r[0] != REC_NULL and r[0] >= 16:
                alive[i] = (int(r[2]), 1 << (int(r[0]) - 16))
            elif k in (2, 3) and r[1] == LARGE_SB:
                for j, (u, n) in list(alive.items()):
                    if u == int(r[2]):
                        del alive[j]
    return list(alive.values())


def class_switch_program(nsets=32, first=64, then=16, nw=64):
    """One wave per step on head set step % nsets: nw steps of `first` B
    blocks, frees of all of them, then nw steps of `then` B blocks.
    Returns (ops, head set per step)."""
    lanes = 64
    a = np.zeros((nw, lanes, 2), dtype=np.uint32)
    a[:, :, 0] = 1
    a[:, :, 1] = first
    f = np.zeros((nw, lanes, 2), dtype=np.uint32)
    f[:, :, 0] = 2
    f[:, :, 1] = np.arange(nw * lanes).reshape(nw, lanes)
    b = a.copy()
    b[:, :, 1] = then
    ops = np.concatenate([a, f, b])
    return ops, (np.arange(3 * nw) % nsets).astype(np.uint32)


def head_words(dump, nsbs, nheads):
    """(heads, cached heads) of a state dump: nheads x 16 slab words."""
    from oracle.model import NSETS
    hb = 3 * nsbs + NSETS * ((nsbs + 63) // 64)
    n = nheads * 16
    return dump[hb:hb + n].reshape(nheads, 16), dump[hb + n:hb + 2 * n].reshape(nheads, 16)
