"""The five benchmark workloads: seeded inputs and query batches.

Every input comes from the library's splitmix64 host generators and every
query batch from numpy's PCG64 (`default_rng`) with a fixed seed, so the bench,
the GPU tests and the scripts that write the golden digests all regenerate the
same data bit for bit.

  c1  levelWT bytes: n = 2^24 Zipf(1.1) bytes over a seeded byte permutation,
      8 levels; 10^6 rank(c, i) queries
  c2  wt_int: n = 2^28 uint32, uniform in [0, 2^20) and geometric(1e-4) capped
      at 2^20 - 1, max_level = 20; 10^7 rank1(i) and 10^7 select1(j) queries
      over the concatenated tree bits
  c3  FM-index: 2^27 bytes of order-5 Markov DNA (text[n-1] = 0); 10^6
      length-20 patterns (even rows exact, odd rows 1-2 substitutions)
  c4  batched trees: 4096 byte sequences, n_i = floor(2^U(10,22)), alphabets
      {4, 20, 256}
  c5  learned select: 64 bitvectors of 2^26 bits (densities log-uniform in
      [0.01, 0.5], even Bernoulli, odd run-clustered); 10^8 uniform
      (vector, rank) select1 queries
"""
import numpy as np

SEED = 0x5EED2026


def _pkg():
    import sys
    return sys.modules[__name__.rsplit(".", 1)[0]]


# ------------------------------------------------------------------ c1
C1_N, C1_LEVELS, C1_QUERIES = 1 << 24, 8, 1_000_000


def c1_input():
    return _pkg().gen_zipf_bytes(SEED + 1, C1_N, 1.1)


def c1_queries(n=C1_N, nq=C1_QUERIES):
    """(i, c): i uniform in [0, n], c uniform in [0, 256)"""
    rng = np.random.default_rng(SEED + 2)
    qi = rng.integers(0, n + 1, nq).astype(np.uint64)
    qc = rng.integers(0, 256, nq).astype(np.uint64)
    return qi, qc


# ------------------------------------------------------------------ c2
C2_N, C2_LEVELS, C2_QUERIES = 1 << 28, 20, 10_000_000


def c2_input(kind="uniform"):
    p = _pkg()
    if kind == "uniform":
        return p.gen_uniform(SEED, C2_N, 1 << C2_LEVELS, np.uint32)
    return p.gen_geometric(SEED + 9, C2_N, 1e-4, (1 << C2_LEVELS) - 1, np.uint32)


def c2_queries(nbits, ones, nq=C2_QUERIES):
    """(rank positions in [0, nbits], select1 ranks in [1, ones])"""
    rng = np.random.default_rng(SEED + 1)
    return (rng.integers(0, nbits + 1, nq, dtype=np.uint64),
            rng.integers(1, ones + 1, nq, dtype=np.uint64))


# ------------------------------------------------------------------ c3
C3_N, C3_M, C3_PATTERNS = 1 << 27, 20, 1_000_000


def c3_text(n=C3_N):
    return _pkg().gen_dna(SEED + 3, n)


def c3_patterns(text, npat=C3_PATTERNS):
    return _pkg().gen_patterns(SEED + 4, text, C3_M, npat)


# ------------------------------------------------------------------ c4
C4_SEQS = 4096


def c4_batch(nseq=C4_SEQS, log2_max=22.0, symbols=True):
    """(n[nseq], sigma[nseq], symbols, offsets)"""
    n, sigma, sym = _pkg().gen_batch(SEED + 7, nseq, 10.0, float(log2_max), symbols=symbols)
    off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)
    return n, sigma, sym, off


# ------------------------------------------------------------------ c5
C5_VECS, C5_LOG2BITS, C5_QUERIES = 64, 26, 100_000_000


def c5_vectors(nvec=C5_VECS, log2bits=C5_LOG2BITS):
    """(words[nvec, nbits/64], density[nvec])"""
    return _pkg().gen_select_vectors(SEED + 5, nvec, 1 << log2bits)


def c5_queries(ones, nq=C5_QUERIES):
    """(vector ids uint32, 1-based ranks uint64): vector uniform, rank uniform in
    [1, ones[vector]]"""
    ones = np.asarray(ones, np.uint64)
    rng = np.random.default_rng(SEED + 6)
    vec = rng.integers(0, ones.size, nq).astype(np.uint32)
    j = (rng.random(nq) * ones[vec]).astype(np.uint64) + 1
    return vec, j
