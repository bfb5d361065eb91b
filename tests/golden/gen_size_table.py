"""Writes size_table_ref_sb22.json: the reference allocator's size table at
4 MiB slabs (sb_sz_sh 22) with its default fractions (busy .835, roomy .6,
sparse .012; src/halloc.h:30-32).

The reference is CUDA and cannot run here, so this script restates, in plain
Python, the host code that builds the table: src/halloc.cu:363-394 (block
size, nblocks rounded to tmod_by_size of src/halloc.cu:101-104, chunk id and
size, blocks per chunk, hash step, truncated thresholds) and max_prime_below
of src/utils.cu:8-24 (a prime found by a sieve here, not by trial division).
It shares no code with the host model it checks (oracle/halloc_model.hpp).

    python tests/golden/gen_size_table.py        # rewrites the JSON
    python tests/golden/gen_size_table.py --check # exits 1 if it differs
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "size_table_ref_sb22.json")
SOURCE = ("reference size table (src/halloc.cu:367-393) at sb_sz_sh=22, busy .835, roomy .6, sparse .012; "
          "written by tests/golden/gen_size_table.py")


def primes_upto(n):
    sieve = bytearray([1]) * (n + 1)
    sieve[0:2] = b"\x00\x00"
    for i in range(2, int(n ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(sieve[i * i::i]))
    return [i for i in range(n + 1) if sieve[i]]


def max_prime_below(n, nb):
    """The largest prime p >= 3 below n with n % p and nb % p nonzero."""
    for p in reversed(primes_upto(max(n - 1, 2))):
        if p >= 3 and n % p and nb % p:
            return p
    raise ValueError("no prime")


def table(sb_sz_sh=22, busy=0.835, roomy=0.6, sparse=0.012):
    sb_sz = 1 << sb_sz_sh
    rows = []
    for isize in range(16):
        unit = 1 << (isize // 2 + 3)
        block_sz = 3 * unit if isize % 2 else 2 * unit
        tmod = 1 << max(2 - isize // 2, 0)          # THREAD_MOD_SH = 2
        nblocks = sb_sz // block_sz // tmod * tmod
        chunk_id = isize % 2 + (0 if isize < 8 else 2)
        chunk_bytes = (3 if chunk_id % 2 else 2) * (128 if chunk_id // 2 else 8)
        ncib = block_sz // chunk_bytes
        nchunks = nblocks * ncib
        rows.append({
            "id": isize, "block_sz": block_sz, "chunk_id": chunk_id, "chunk_bytes": chunk_bytes, "ncib": ncib,
            "nblocks": nblocks, "nchunks": nchunks,
            "hash_step": ncib * max_prime_below(nblocks // 256 + nblocks // 64, nblocks),
            "roomy": int(roomy * nchunks), "busy": int(busy * nchunks), "sparse": int(sparse * nchunks),
        })
    return rows


def main():
    doc = {"source": SOURCE, "rows": table()}
    text = json.dumps(doc, indent=1) + "\n"
    if "--check" in sys.argv:
        old = json.load(open(OUT))
        sys.exit(0 if old["rows"] == doc["rows"] else 1)
    with open(OUT, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main()
