"""Synthetic protein workloads for tests and benchmarks.

Residues are drawn i.i.d. from the Robinson & Robinson (1991, PNAS 88:8880)
background amino-acid frequencies; lengths are lognormal (mean 350, sd 250 by
default) clipped to a range; homologs are derived from a source sequence by
BLOSUM62-conditional substitution at a target identity plus insertions and
deletions at a per-residue rate with geometric lengths.  Everything is
vectorised numpy and seeded, so a (seed, sizes) pair names a workload.
"""
import numpy as np

from .matrices import MATRIX_ORDER, matrix

# Robinson & Robinson background frequencies, per mille.
RR_PER_MILLE = {
    "A": 78.05, "C": 19.25, "D": 53.64, "E": 62.95, "F": 38.56, "G": 73.77,
    "H": 21.99, "I": 51.42, "K": 57.44, "L": 90.19, "M": 22.43, "N": 44.87,
    "P": 52.03, "Q": 42.64, "R": 51.29, "S": 71.20, "T": 58.41, "V": 64.41,
    "W": 13.30, "Y": 32.16,
}


def background():
    """Probabilities over the 25 matrix indices (0 for B, J, Z, X, *)."""
    p = np.zeros(25)
    for aa, f in RR_PER_MILLE.items():
        p[MATRIX_ORDER.index(aa)] = f
    return p / p.sum()


def lognormal_lengths(rng, n, mean=350.0, sd=250.0, lo=30, hi=5000):
    sigma2 = np.log1p((sd / mean) ** 2)
    mu = np.log(mean) - sigma2 / 2
    x = rng.lognormal(mu, np.sqrt(sigma2), n)
    return np.clip(np.rint(x), lo, hi).astype(np.int32)


def _quantile_table(cdf):
    """Map a 16-bit uniform draw to a residue: entry u is the first index
    whose cumulative probability exceeds (u + 0.5) / 65536."""
    u = (np.arange(65536) + 0.5) / 65536
    return np.minimum(np.searchsorted(cdf, u, side="right"), len(cdf) - 1).astype(np.int8)


_BG_TABLE = None


def random_residues(rng, total):
    global _BG_TABLE
    if _BG_TABLE is None:
        _BG_TABLE = _quantile_table(np.cumsum(background()))
    return _BG_TABLE[rng.integers(0, 65536, int(total), dtype=np.uint16)]


def concat_random(rng, lengths):
    """(flat residues, offsets, lengths) for i.i.d. background sequences."""
    lengths = np.asarray(lengths, dtype=np.int32)
    off = np.zeros(len(lengths), np.int64)
    off[1:] = np.cumsum(lengths[:-1], dtype=np.int64)
    return random_residues(rng, lengths.sum()), off, lengths


def _substitution_cdf():
    """Row a: cumulative P(b | a, b != a) ~ p_b exp(lambda_u s62(a, b)) with
    lambda_u = ln 2 / 2, over standard residues."""
    p = background()
    s = matrix("BLOSUM62").reshape(25, 25).astype(np.float64)
    w = p[None, :] * np.exp(np.log(2) / 2 * s)
    np.fill_diagonal(w, 0.0)
    w[:, p == 0] = 0.0
    w /= np.maximum(w.sum(axis=1, keepdims=True), 1e-300)
    return np.cumsum(w, axis=1)


_CDF = None
_CDF_FLAT = None


def mutate(rng, res, off, lens, identity, indel_rate=0.02, ext_p=0.5):
    """Homologs of the given sequences.

    res/off/lens: flat source sequences; identity: per-sequence target identity.
    Returns (flat residues, offsets, lengths) of the homologs.
    """
    global _CDF, _CDF_FLAT
    if _CDF is None:
        _CDF = _substitution_cdf()
        _CDF_FLAT = np.concatenate([_quantile_table(row) for row in _CDF])
    lens = np.asarray(lens, np.int64)
    n = len(lens)
    seg = np.repeat(np.arange(n), lens)
    total = int(lens.sum())
    starts_out = np.zeros(n, np.int64)
    if n:
        starts_out[1:] = np.cumsum(lens)[:-1]
    gather = np.repeat(np.asarray(off, np.int64) - starts_out, lens) + np.arange(total)
    src = res[gather].astype(np.int64)
    ident = np.asarray(identity, np.float64)[seg]
    # substitutions
    sub = rng.random(len(src)) >= ident
    u = rng.integers(0, 65536, int(sub.sum()), dtype=np.uint16)
    a = src[sub]
    b = _CDF_FLAT[a * 65536 + u]
    out = src.copy()
    out[sub] = b
    # indel events: half deletions, half insertions, geometric lengths
    ev = rng.random(len(out)) < indel_rate
    is_del = ev & (rng.random(len(out)) < 0.5)
    is_ins = ev & ~is_del
    starts = np.zeros(n, np.int64)
    starts[1:] = np.cumsum(lens)[:-1]
    ends = starts + lens
    # deletion runs, clipped at the end of their sequence (difference array)
    dpos = np.nonzero(is_del)[0]
    dend = np.minimum(dpos + rng.geometric(ext_p, len(dpos)), ends[seg[dpos]])
    diff = (np.bincount(dpos, minlength=len(out) + 1) -
            np.bincount(dend, minlength=len(out) + 1))
    keep = np.cumsum(diff[:-1]) == 0
    keep[starts[lens > 0]] = True       # never delete a whole sequence
    ins_len = np.zeros(len(out), np.int64)
    ins_len[is_ins] = rng.geometric(ext_p, int(is_ins.sum()))
    rep = keep.astype(np.int64) + ins_len
    total = int(rep.sum())
    new = np.empty(total, np.int64)
    first = np.zeros(len(out), np.int64)
    first[1:] = np.cumsum(rep)[:-1]
    # each kept source residue is followed by its insertion run
    new[:] = -1
    kept_idx = np.nonzero(keep)[0]
    new[first[kept_idx]] = out[kept_idx]
    fill = new < 0
    new[fill] = random_residues(rng, int(fill.sum()))
    new_seg = np.repeat(seg, rep)
    new_lens = np.bincount(new_seg, minlength=n).astype(np.int32)
    new_off = np.zeros(n, np.int64)
    new_off[1:] = np.cumsum(new_lens[:-1], dtype=np.int64)
    return new.astype(np.int8), new_off, new_lens


def pair_workload(rng, n_queries, subjects_per_query, homolog_frac=0.5,
                  ident_range=(0.25, 0.95), mean=350.0, sd=250.0, lo=30, hi=5000):
    """Queries, subjects and pairs for the batched-SW kernel benchmark:
    every query is paired with subjects_per_query subjects of its own, a
    fraction homolog_frac of them homologs of the query.

    Returns dict with q=(res, off, len), s=(res, off, len), pair_query,
    pair_subject, is_homolog.
    """
    ql = lognormal_lengths(rng, n_queries, mean, sd, lo, hi)
    q = concat_random(rng, ql)
    n_hom = int(round(subjects_per_query * homolog_frac))
    n_rand = subjects_per_query - n_hom
    # homologs of every query
    rep_q = np.repeat(np.arange(n_queries), n_hom)
    ident = rng.uniform(ident_range[0], ident_range[1], len(rep_q))
    h_res, h_off, h_len = mutate(rng, q[0], q[1][rep_q], q[2][rep_q], ident)
    h_len_c = np.clip(h_len, 1, None)
    # random subjects
    rl = lognormal_lengths(rng, n_queries * n_rand, mean, sd, lo, hi)
    r_res, r_off, r_len = concat_random(rng, rl)
    # subjects: homologs then randoms
    s_res = np.concatenate([h_res, r_res])
    s_off = np.concatenate([h_off, r_off + len(h_res)])
    s_len = np.concatenate([h_len, r_len]).astype(np.int32)
    # clip homolog lengths into range by truncation (and pad the rare empty one)
    too_long = s_len > hi
    s_len[too_long] = hi
    pair_query = np.concatenate([rep_q, np.repeat(np.arange(n_queries), n_rand)]).astype(np.int32)
    pair_subject = np.arange(len(s_len), dtype=np.int32)
    is_hom = np.concatenate([np.ones(len(rep_q), bool), np.zeros(len(rl), bool)])
    del h_len_c
    return dict(q=q, s=(s_res, s_off, s_len), pair_query=pair_query,
                pair_subject=pair_subject, is_homolog=is_hom)


def to_letters(res):
    """Matrix indices -> amino-acid letters (substitution.h order)."""
    table = np.frombuffer(MATRIX_ORDER.encode(), np.uint8)
    return table[np.asarray(res, np.int64)].tobytes().decode()


def write_fasta(path, seqs, prefix="seq"):
    """seqs: iterable of matrix-index arrays; 60 letters per line."""
    with open(path, "w") as f:
        for k, s in enumerate(seqs):
            f.write(">%s%d\n" % (prefix, k))
            letters = to_letters(s)
            for i in range(0, len(letters), 60):
                f.write(letters[i:i + 60] + "\n")


def search_database(rng, n_db, n_queries, hom_frac=0.2, max_hom=5, ident=(0.3, 0.9),
                    mean=350.0, sd=250.0, lo=30, hi=2000):
    """A database with planted homologs of a fraction of the queries (the
    end-to-end search workload): returns (db_seqs, query_seqs)."""
    ql = lognormal_lengths(rng, n_queries, mean, sd, lo, hi)
    qr, qo, qn = concat_random(rng, ql)
    queries = [qr[o:o + n] for o, n in zip(qo, qn)]
    homs = []
    for k in range(n_queries):
        if rng.random() < hom_frac:
            for _ in range(int(rng.integers(1, max_hom + 1))):
                r_, o_, l_ = mutate(rng, queries[k], np.array([0]), np.array([len(queries[k])]),
                                    np.array([rng.uniform(*ident)]))
                homs.append(r_[:l_[0]])
    n_rand = max(0, n_db - len(homs))
    dl = lognormal_lengths(rng, n_rand, mean, sd, lo, hi)
    dr, do, dn = concat_random(rng, dl)
    db = [dr[o:o + n] for o, n in zip(do, dn)] + homs
    order = rng.permutation(len(db))
    return [db[i] for i in order], queries


def internal_text_and_seeds(rng, t_residues, n_queries, n_seeds, homolog_frac=0.05,
                            mean=350.0, sd=250.0, lo=30, hi=5000):
    """The gapless-extension workload: a newline-separated internal-alphabet
    text of about t_residues residues (lognormal protein lengths), n_queries
    queries, and n_seeds seeds (query, offset, text position) drawn uniformly
    over query x text positions, with a fraction homolog_frac placed on the
    diagonals of homolog copies of the queries planted in the text.

    Returns (text bytes as uint8 array, list of query byte strings, seeds as a
    structured array with fields query/offset/position, newline positions).
    """
    internal = np.zeros(25, np.uint8)
    from .matrices import INTERNAL_ORDER
    for k, aa in enumerate(INTERNAL_ORDER):
        internal[MATRIX_ORDER.index(aa)] = ord("A") + k
    ql = lognormal_lengths(rng, n_queries, mean, sd, lo, hi)
    qr, qo, qn = concat_random(rng, ql)
    queries = [internal[qr[o:o + n]].tobytes() for o, n in zip(qo, qn)]
    # planted homologs: one mutated copy of each query (first n_plant)
    n_plant = min(n_queries, 2000)
    hr, ho, hn = mutate(rng, qr, qo[:n_plant], qn[:n_plant], rng.uniform(0.3, 0.9, n_plant))
    n_prot = int(t_residues / mean)
    lens = lognormal_lengths(rng, n_prot, mean, sd, lo, hi)
    # text: proteins with '\n' after each; the planted copies go first
    all_lens = np.concatenate([hn, lens]).astype(np.int64)
    total = int(all_lens.sum() + len(all_lens))
    text = np.empty(total, np.uint8)
    starts = np.zeros(len(all_lens), np.int64)
    starts[1:] = np.cumsum(all_lens[:-1] + 1)
    nl = starts + all_lens
    body = internal[random_residues(rng, total)]
    text[:] = body
    text[nl] = ord("\n")
    # copy the planted homologs in (they occupy the first n_plant proteins)
    first = np.concatenate([[0], np.cumsum(hn[:-1])]).astype(np.int64)
    pos_in_text = np.repeat(starts[:n_plant] - first, hn) + np.arange(int(hn.sum()))
    text[pos_in_text] = internal[hr]
    seeds = np.zeros(n_seeds, dtype=[("query", "<i4"), ("offset", "<i4"), ("position", "<i8")])
    n_hom = int(n_seeds * homolog_frac)
    q = rng.integers(0, n_queries, n_seeds)
    seeds["query"] = q
    seeds["offset"] = (rng.random(n_seeds) * ql[q]).astype(np.int32)
    seeds["position"] = rng.integers(0, total, n_seeds)
    # homolog seeds: a planted copy's start plus the offset (its diagonal)
    hq = rng.integers(0, n_plant, n_hom)
    off = (rng.random(n_hom) * np.minimum(ql[hq], hn[hq])).astype(np.int64)
    seeds["query"][:n_hom] = hq
    seeds["offset"][:n_hom] = off
    seeds["position"][:n_hom] = starts[hq] + off
    return text, queries, seeds, nl


def matching_seeds(rng, queries, text, n_seeds, homolog_seeds=None):
    """Seeds whose first residues agree, as suffix-array seeds always do: a
    text position uniform over the residues of the text, then a query
    position uniform over the query residues equal to it.  homolog_seeds
    (structured seeds, e.g. the planted-diagonal ones of
    internal_text_and_seeds) fill the front of the array where their first
    residues agree."""
    ql = np.array([len(q) for q in queries], np.int64)
    qoff = np.zeros(len(ql), np.int64)
    qoff[1:] = np.cumsum(ql[:-1])
    flat = np.frombuffer(b"".join(queries), np.uint8)
    order = np.argsort(flat, kind="stable")
    counts = np.bincount(flat, minlength=256)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    qid = np.repeat(np.arange(len(ql)), ql)
    seeds = np.zeros(n_seeds, dtype=[("query", "<i4"), ("offset", "<i4"), ("position", "<i8")])
    k0 = 0
    if homolog_seeds is not None and len(homolog_seeds):
        h = homolog_seeds
        ok = text[h["position"]] == flat[qoff[h["query"]] + h["offset"]]
        h = h[ok][:n_seeds]
        seeds[:len(h)] = h
        k0 = len(h)
    pos = rng.integers(0, len(text), n_seeds - k0)
    while True:
        bad = np.nonzero(counts[text[pos]] == 0)[0]     # '\n' or a letter no query has
        if not len(bad):
            break
        pos[bad] = rng.integers(0, len(text), len(bad))
    c = text[pos]
    k = first[c] + (rng.random(len(pos)) * counts[c]).astype(np.int64)
    j = order[k]
    seeds["query"][k0:] = qid[j]
    seeds["offset"][k0:] = j - qoff[qid[j]]
    seeds["position"][k0:] = pos
    return seeds


def newline_rank(text):
    """rank[i] = number of '\n' in text[0..i] (int32): with it the first
    newline at or after a position is found by one gather instead of a
    binary search."""
    return np.cumsum(text == 10, dtype=np.int32)


def gapless_scan_bytes(seeds, q_len, text, nl, rank=None):
    """Text bytes the reference rule reads per seed (search.c:99-190): the
    first residue when it differs from the query's, else the right window up
    to and including a '\\n' (at most min(Tsize-idx, m-o, 128)) and the left
    window i = 1 .. min(idx, o, 128)-1 up to and including a '\\n'."""
    idx = seeds["position"].astype(np.int64)
    o = seeds["offset"].astype(np.int64)
    m = np.asarray(q_len, np.int64)[seeds["query"]]
    tsize = len(text)
    if rank is None:
        k = np.searchsorted(nl, idx)             # first newline at or after idx
    else:
        k = rank[idx].astype(np.int64) - (text[idx] == 10)
    nxt = np.where(k < len(nl), nl[np.minimum(k, len(nl) - 1)], tsize)
    prv = np.where(k > 0, nl[np.maximum(k - 1, 0)], -1)
    R = np.minimum(np.minimum(tsize - idx, m - o), 128)
    right = np.minimum(R, nxt - idx + 1)
    Lw = np.minimum(np.minimum(idx, o), 128)
    left = np.maximum(np.minimum(Lw - 1, idx - prv), 0)
    return right, left


_LETTERS = np.frombuffer(MATRIX_ORDER.encode(), np.uint8)


def _write_fasta_flat(f, res, lens, names_from, prefix, chunk=1 << 26):
    """FASTA (60 letters per line) of sequences stored back to back in `res`
    with lengths `lens`, written in chunks of about `chunk` residues with
    numpy: sequence k is named prefix + str(names_from + k)."""
    lens = np.asarray(lens, np.int64)
    ends = np.cumsum(lens)
    k0 = 0
    while k0 < len(lens):
        r0 = int(ends[k0 - 1]) if k0 else 0
        k1 = max(k0 + 1, int(np.searchsorted(ends, r0 + chunk, side="right")))
        k1 = min(k1, len(lens))
        L = lens[k0:k1]
        n_res = int(L.sum())
        heads = "".join(">%s%d\n" % (prefix, names_from + k) for k in range(k0, k1)).encode()
        hl = np.array([len(">%s%d\n" % (prefix, names_from + k)) for k in range(k0, k1)],
                      np.int64)
        nlines = (L + 59) // 60
        rec = hl + L + nlines
        base = np.zeros(len(L), np.int64)
        base[1:] = np.cumsum(rec[:-1])
        out = np.full(int(rec.sum()), ord("\n"), np.uint8)
        # residue g of the chunk goes to g + (bytes of headers and line ends
        # before it): a prefix sum over a sparse increment array
        start = np.zeros(len(L), np.int64)
        start[1:] = np.cumsum(L[:-1])
        inc = np.zeros(n_res + 1, np.int32)
        inc[start] = ((base - start) + hl - np.concatenate(
            [[0], base[:-1] - start[:-1] + hl[:-1] + (L[:-1] - 1) // 60])).astype(np.int32)
        # one byte more at every line start after the first of a sequence
        extra_lines = nlines - 1
        line_seq = np.repeat(np.arange(len(L)), extra_lines)
        t = np.arange(int(extra_lines.sum())) - np.repeat(np.cumsum(extra_lines) - extra_lines,
                                                          extra_lines)
        inc[start[line_seq] + 60 * (t + 1)] = 1       # (line starts never hit a sequence start)
        dest = np.cumsum(inc[:n_res], dtype=np.int64)
        dest += np.arange(n_res, dtype=np.int64)
        out[dest] = _LETTERS[res[r0:r0 + n_res].astype(np.int64)]
        # headers
        hpos = np.repeat(base, hl) + (np.arange(int(hl.sum())) -
                                       np.repeat(np.cumsum(hl) - hl, hl))
        out[hpos] = np.frombuffer(heads, np.uint8)
        f.write(out.tobytes())
        k0 = k1


def write_search_fasta(db_path, q_path, rng, n_db, n_queries, hom_frac=0.1, max_hom=5,
                       ident=(0.3, 0.9), mean=350.0, sd=250.0, lo=30, hi=2000):
    """The end-to-end search workload written straight to FASTA, vectorised
    for databases of millions of proteins: n_queries i.i.d. queries, a
    fraction hom_frac of them with U{1..max_hom} planted homologs (identity
    uniform in `ident`, 2 % indels) placed at random database positions, the
    other database proteins i.i.d.  Returns (database residues, queries
    residues) as (flat, offsets, lengths) triples."""
    ql = lognormal_lengths(rng, n_queries, mean, sd, lo, hi)
    qr, qo, qn = concat_random(rng, ql)
    bearing = np.nonzero(rng.random(n_queries) < hom_frac)[0]
    reps = rng.integers(1, max_hom + 1, len(bearing))
    src = np.repeat(bearing, reps)
    hr, ho, hn = mutate(rng, qr, qo[src], qn[src], rng.uniform(ident[0], ident[1], len(src)))
    n = max(n_db, len(src))
    lens = lognormal_lengths(rng, n, mean, sd, lo, hi).astype(np.int64)
    slots = rng.permutation(n)[:len(src)]
    lens[slots] = hn
    off = np.zeros(n, np.int64)
    off[1:] = np.cumsum(lens[:-1])
    res = random_residues(rng, int(lens.sum()))
    # the homologs' residues into their slots
    seg = np.repeat(np.arange(len(src)), hn)
    j = np.arange(int(hn.sum())) - np.repeat(np.cumsum(hn) - hn, hn)
    res[np.repeat(off[slots], hn) + j] = hr[np.repeat(ho, hn) + j]
    del seg
    with open(db_path, "wb") as f:
        _write_fasta_flat(f, res, lens, 0, "s")
    with open(q_path, "wb") as f:
        _write_fasta_flat(f, qr, qn, 0, "q")
    return (res, off, lens), (qr, qo, qn)


def search_text(rng, n_db, n_queries, hom_frac=0.1, max_hom=5, ident=(0.3, 0.9), mean=350.0,
                sd=250.0, lo=30, hi=2000):
    """The end-to-end search workload as the index's text directly (for
    databases of millions of proteins): the .psq text of n_db proteins
    (internal letters, '\n' after each; i.i.d. background residues, with
    U{1..max_hom} planted homologs of a fraction hom_frac of the queries at
    random positions, identity uniform in `ident`, 2 % indels), their
    lengths, and the queries as (flat residues, offsets, lengths)."""
    from .index import INTERNAL_OF_INDEX
    global _BG_TABLE
    ql = lognormal_lengths(rng, n_queries, mean, sd, lo, hi)
    qr, qo, qn = concat_random(rng, ql)
    bearing = np.nonzero(rng.random(n_queries) < hom_frac)[0]
    reps = rng.integers(1, max_hom + 1, len(bearing))
    src = np.repeat(bearing, reps)
    hr, ho, hn = mutate(rng, qr, qo[src], qn[src], rng.uniform(ident[0], ident[1], len(src)))
    n = max(n_db, len(src))
    lens = lognormal_lengths(rng, n, mean, sd, lo, hi).astype(np.int64)
    slots = rng.permutation(n)[:len(src)]
    lens[slots] = hn
    starts = np.cumsum(lens + 1) - (lens + 1)
    if _BG_TABLE is None:
        _BG_TABLE = _quantile_table(np.cumsum(background()))
    letters = INTERNAL_OF_INDEX[_BG_TABLE.astype(np.int64)]
    total = int(lens.sum()) + n
    text = np.empty(total, np.uint8)
    step = 1 << 28
    for a in range(0, total, step):
        b = min(total, a + step)
        text[a:b] = letters[rng.integers(0, 65536, b - a, dtype=np.uint16)]
    text[starts + lens] = ord("\n")
    j = np.arange(int(hn.sum())) - np.repeat(np.cumsum(hn) - hn, hn)
    text[np.repeat(starts[slots], hn) + j] = INTERNAL_OF_INDEX[hr[np.repeat(ho, hn) + j].astype(np.int64)]
    return text, lens, (qr, qo, qn)


def write_queries_fasta(path, queries, prefix="q"):
    """Queries given as (flat residues, offsets, lengths), stored back to
    back, to FASTA."""
    qr, qo, qn = queries
    with open(path, "wb") as f:
        _write_fasta_flat(f, qr, qn, 0, prefix)
