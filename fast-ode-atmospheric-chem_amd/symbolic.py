"""Host symbolic analysis of the stage matrix alpha*M - J for the device LU.

The reference factors alpha*M - J with SuperLU (``permc_spec="MMD_AT_PLUS_A"``,
partial pivoting) on a block-diagonal matrix of the group's cells
(fastode/linalg.py:70-81), or with LAPACK getrf per cell (linalg.py:15-21).
The device factors each cell's matrix in place *without pivoting*, on a fill
pattern fixed here once per mechanism:

1. order: minimum degree on the pattern of A + A^T (ties to the lowest index),
   the same family of ordering SuperLU is asked for;
2. fill: symbolic elimination without pivoting gives the L+U pattern;
3. split: pivots [0, c) are *sparse* and [c, n) form the *core*, a trailing
   block that fill has made nearly dense.  The core is factored as a dense
   block held in registers, distributed 2-D cyclically over a G x G thread
   grid with TILE x TILE entries per thread (m_pad = G * TILE >= n - c,
   padded with identity rows);
4. sparse entries (min(i, j) < c) live in LDS and are computed by
   dot-product tasks  A_ij = (A_ij - sum_{p < min(i,j)} L_ip U_pj) [/ U_jj],
   grouped in stages: a task runs once every task it reads has run;
5. the core's Schur complement update from the sparse pivots,
   S_ij = -sum_{p < c} L_ip U_pj, is formed by the thread that owns the core
   entry, straight into its register tile, from its own list of pairs;
6. the solves use level-scheduled row tasks for the sparse rows and block
   substitution with inverted diagonal blocks for the core.

LDS index space of one cell: [0, ns) sparse entries, then the solve vector
(c + m_pad entries).

Table formats (all little integers, so that they fit in LDS):
* task  = 2 x uint32: [target | divisor << 16] (divisor 0xFFFF: none) and
  [first pair | count << 20];
* pair  = uint32 a << 16 | b  (LDS indices of the two factors);
* core_src = int16 [thread][slot_stride]: Jacobian entry, -1 fill, -2 the
  identity diagonal of a padding row;
* sch_mask = uint32 [thread][4]: bit s set when core slot s has a Schur sum;
  the thread's pairs contiguous in sch_pairs from sch_base[thread], bit 31
  of the pair that closes a sum.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

G = 16                 # core thread grid is G x G (256 threads)
NT = G * G
TILE_MAX = 10          # largest supported core: G * TILE_MAX = 160
DENSE_RHO = 0.85       # core = longest trailing block at least this dense
NO_DIV = 0xFFFF


def min_degree_order(n, rows, cols):
    """Minimum degree ordering of the graph of A + A^T (lowest index on ties).
    Returns perm with perm[new] = old."""
    Gm = np.zeros((n, n), dtype=bool)
    Gm[rows, cols] = True
    Gm |= Gm.T
    np.fill_diagonal(Gm, False)
    alive = np.ones(n, dtype=bool)
    order = []
    for _ in range(n):
        deg = np.where(alive, Gm.sum(axis=1), np.iinfo(np.int64).max)
        v = int(np.argmin(deg))
        order.append(v)
        nb = np.nonzero(Gm[v] & alive)[0]
        Gm[np.ix_(nb, nb)] = True
        Gm[nb, nb] = False
        alive[v] = False
        Gm[v, :] = False
        Gm[:, v] = False
    return np.asarray(order, dtype=np.int64)


def symbolic_fill(n, rows_new, cols_new):
    """Filled L+U pattern of no-pivot elimination, as a dense bool matrix."""
    F = np.zeros((n, n), dtype=bool)
    F[rows_new, cols_new] = True
    F[np.arange(n), np.arange(n)] = True
    for k in range(n):
        L = np.nonzero(F[k + 1:, k])[0] + k + 1
        U = np.nonzero(F[k, k + 1:])[0] + k + 1
        if len(L) and len(U):
            F[np.ix_(L, U)] = True
    return F


def choose_split(F, rho=DENSE_RHO, m_max=G * TILE_MAX):
    """Core start c: the longest trailing block at least rho dense, no larger
    than m_max, and at least one pivot."""
    n = F.shape[0]
    c = n - 1
    for cand in range(n - 1, -1, -1):
        if F[cand:, cand:].mean() >= rho:
            c = cand
    return int(max(c, n - m_max))


def _stage(levels, key=None):
    """Sort task ids by level (then by ``key``, stably); return (order,
    stage_ptr)."""
    levels = np.asarray(levels, dtype=np.int64)
    if len(levels) == 0:
        return np.zeros(0, dtype=np.int64), np.zeros(1, dtype=np.int64)
    if key is None:
        order = np.argsort(levels, kind="stable")
    else:
        order = np.lexsort((np.asarray(key, dtype=np.int64), levels))
    nst = int(levels.max()) + 1
    counts = np.bincount(levels, minlength=nst)
    return order, np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)


def _pack_tasks(tgt, start, count, div):
    k = len(tgt)
    out = np.zeros((k, 2), dtype=np.uint32)
    if k:
        tgt, start, count, div = (np.asarray(x, dtype=np.int64) for x in (tgt, start, count, div))
        if (start >= 1 << 20).any() or (count >= 1 << 12).any():
            raise ValueError("task pair ranges exceed the packed format")
        out[:, 0] = tgt | (np.where(div < 0, NO_DIV, div) << 16)
        out[:, 1] = start | (count << 20)
    return out.ravel()


def _pack_pairs(a, b):
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    return ((a << 16) | b).astype(np.uint32)


@dataclass
class LUPlan:
    n: int
    nnz: int                 # Jacobian pattern entries
    perm: np.ndarray         # new -> old
    c: int                   # first core pivot
    m: int                   # core size n - c
    tile: int                # core tile per thread
    F: np.ndarray            # filled pattern (new ordering)
    ns: int                  # sparse LU entries (stored in LDS)
    nsch: int                # core entries with a Schur sum
    nnz_lu: int              # all L+U entries of the filled pattern
    arrays: dict             # flat arrays uploaded to the device

    @property
    def m_pad(self):
        return G * self.tile

    @property
    def slot_stride(self):
        return -(-self.tile * self.tile // 8) * 8

    @property
    def lds_vec(self):
        return self.ns

    @property
    def lds_len(self):
        return self.ns + self.c + self.m_pad

    @property
    def sch_depth(self):
        """Most Schur sums one thread holds."""
        m = self.arrays["sch_mask"].reshape(NT, 4)
        return int(max(sum(bin(int(x)).count("1") for x in row) for row in m))

    @property
    def lu_len(self):
        """Doubles per cell of a stored factorization: sparse entries, the core
        tiles and the inverted diagonal blocks."""
        return self.ns + self.m_pad * self.m_pad + 2 * self.tile * NT

    def table_words(self):
        """uint32 words of the tables a persistent workgroup keeps in LDS."""
        a = self.arrays
        return sum(len(a[k]) for k in ("f_tasks", "f_ptr", "f_pairs", "sf_tasks", "sf_ptr", "sb_tasks",
                                        "sb_ptr", "s_pairs", "sch_pairs"))

    def flops(self):
        """Multiply-adds (+ divisions) of one factorization and the L+U entry
        count, as a no-pivot sparse LU would do them."""
        n, F = self.n, self.F
        fact = 0
        for k in range(n):
            nl = int(F[k + 1:, k].sum())
            nu = int(F[k, k + 1:].sum())
            fact += nl * nu + nl
        return fact, int(F.sum())


def build_plan(n, rows, cols, tile=None, rho=DENSE_RHO, order=None, m_max=None):
    """Symbolic analysis of the pattern (rows, cols) (original numbering,
    diagonal included).  ``order`` overrides the ordering (perm[new] = old);
    ``m_max`` caps the core size (a smaller register tile)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    nnz = len(rows)
    if nnz >= 1 << 15:
        raise ValueError("Jacobian patterns above 32767 entries are not supported")
    perm = min_degree_order(n, rows, cols) if order is None else np.asarray(order, dtype=np.int64)
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    rn, cn = inv[rows], inv[cols]
    F = symbolic_fill(n, rn, cn)
    c = choose_split(F, rho, G * TILE_MAX if m_max is None else m_max)
    m = n - c
    if tile is None:
        tile = max(1, -(-m // G))
    if G * tile < m:
        raise ValueError(f"core of {m} rows does not fit tile {tile}")
    if tile > TILE_MAX:
        raise ValueError(f"tile {tile} above the supported {TILE_MAX}")
    m_pad = G * tile

    # ---- sparse storage: entries with min(i, j) < c, row-major
    sp_index = -np.ones((n, n), dtype=np.int64)
    si, sj = np.nonzero(F)
    keep = np.minimum(si, sj) < c
    si, sj = si[keep], sj[keep]
    ns = len(si)
    sp_index[si, sj] = np.arange(ns)

    # ---- core entries updated by sparse pivots get an LDS Schur sum each
    sch_entries = []                            # (ci, cj, [(L_ip, U_pj) ...])
    for ci in range(m):
        i = c + ci
        prow = [p for p in range(c) if F[i, p]]
        if not prow:
            continue
        for cj in range(m):
            j = c + cj
            ps = [p for p in prow if F[p, j]]
            if ps:
                sch_entries.append((ci, cj, [(int(sp_index[i, p]), int(sp_index[p, j])) for p in ps]))
    nsch = len(sch_entries)
    VEC = ns
    if VEC + c + m_pad >= NO_DIV:
        raise ValueError("the LDS index space of one cell exceeds 16 bits")

    # ---- assembly sources.  Entry value = -J[src] (+ alpha*mass[row] on the
    # diagonal).
    jac_of = -np.ones((n, n), dtype=np.int64)
    jac_of[rn, cn] = np.arange(nnz)
    sp_src = jac_of[si, sj].astype(np.int32)
    sp_diag = np.where(si == sj, perm[si], -1).astype(np.int32)     # old row for mass
    sp_row = perm[si].astype(np.int32)                               # old row (alg-only mask)

    stride = -(-tile * tile // 8) * 8
    core_src = np.full((NT, stride), -1, dtype=np.int16)
    core_row = np.full(m_pad, -1, dtype=np.int32)
    core_row[:m] = perm[c:]
    for ti in range(tile):
        for tj in range(tile):
            s = ti * tile + tj
            for tr in range(G):
                for tc in range(G):
                    t = tr * G + tc
                    ci, cj = ti * G + tr, tj * G + tc
                    if ci >= m or cj >= m:
                        if ci == cj:
                            core_src[t, s] = -2
                        continue
                    core_src[t, s] = jac_of[c + ci, c + cj]
    # Schur sums S = 0 - sum_p L_ip U_pj of the core entries, formed by the
    # thread that owns the entry in one pass over its own pairs: its sums in
    # slot order, pairs from sch_base[t] to sch_base[t + 1] (padded with 0 to
    # groups of 4 after the last sum), bit 31 of a pair closing a sum;
    # sch_mask[t] has bit s set when slot s has a sum.
    by_slot = {}
    for ci, cj, prs in sch_entries:
        ti, tr = divmod(ci, G)
        tj, tc = divmod(cj, G)
        by_slot[(tr * G + tc, ti * tile + tj)] = prs
    if by_slot and VEC + c + m_pad >= 1 << 15:
        raise ValueError("Schur pairs need LDS indices below 32768")
    sch_mask = np.zeros((NT, 4), dtype=np.uint32)
    sch_base = np.zeros(NT + 1, dtype=np.int32)
    sch_words = []
    for t in range(NT):
        sch_base[t] = len(sch_words)
        for slot in range(tile * tile):
            prs = by_slot.get((t, slot))
            if prs:
                sch_mask[t, slot >> 5] |= np.uint32(1 << (slot & 31))
                for q, (a_, b_) in enumerate(prs):
                    sch_words.append((int(a_) << 16) | int(b_) | ((1 << 31) if q == len(prs) - 1 else 0))
        while len(sch_words) % 4:
            sch_words.append(0)
    sch_base[NT] = len(sch_words)
    sch_pairs = np.asarray(sch_words, dtype=np.uint32)

    # ---- sparse factorization tasks (dot-product form).  An entry's chain
    # A_ij - L_ip1 U_p1j - L_ip2 U_p2j - ... runs in ascending p, and each
    # multiply-add goes in the first stage after which both of its inputs are
    # final (but not before the previous one's): the chain is cut into
    # consecutive runs, one task per stage, its partial kept in the entry's
    # LDS slot between them, and the division by U_jj goes with the last run.
    # The operations and their order are those of one task in the entry's
    # final stage, so the bits are too; the longest run of a stage, not the
    # longest chain, sets the stage's time.
    tgt, pstart, pcount, div, level = [], [], [], [], []
    pairs_a, pairs_b = [], []
    lev = np.zeros(ns, dtype=np.int64)          # stage after which an entry is final
    # inputs precede outputs: order by min(i, j), then i, j
    order_e = np.lexsort((sj, si, np.minimum(si, sj)))
    for e in order_e:
        i, j = int(si[e]), int(sj[e])
        ps = [p for p in range(min(i, j)) if F[i, p] and F[p, j]]
        d = sp_index[j, j] if i > j else -1
        if not ps and d < 0:
            continue            # the assembled value is final
        start = len(pairs_a)
        ready, at = [], 0
        for p in ps:
            a, b = sp_index[i, p], sp_index[p, j]
            pairs_a.append(a)
            pairs_b.append(b)
            at = max(at, lev[a], lev[b])
            ready.append(at)
        final = max(ready[-1] if ready else 0, lev[d] if d >= 0 else 0)
        q = 0
        while q < len(ps):
            stage = ready[q] if ready[q] != ready[-1] else final
            r = q
            while r < len(ps) and (ready[r] == ready[q]):
                r += 1
            if stage == final:
                r = len(ps)
            tgt.append(e)
            pstart.append(start + q)
            pcount.append(r - q)
            div.append(d if r == len(ps) else -1)
            level.append(stage)
            q = r
        if not ps:
            tgt.append(e)
            pstart.append(start)
            pcount.append(0)
            div.append(d)
            level.append(final)
        lev[e] = final + 1
    fo, fptr = _stage(level, [-c_ for c_ in pcount])
    f_tasks = _pack_tasks(*(np.asarray(x)[fo] if len(tgt) else [] for x in (tgt, pstart, pcount, div)))
    f_pairs = _pack_pairs(pairs_a, pairs_b)

    # ---- solve tasks.  Forward (unit L): v_i -= sum_{p<min(i,c)} L_ip v_p.
    ftgt, fst, fcnt, flev = [], [], [], []
    spa, spb = [], []
    vlev = np.zeros(n, dtype=np.int64)
    for i in range(n):
        ps = [p for p in range(min(i, c)) if F[i, p]]
        if not ps:
            continue
        start = len(spa)
        L = 0
        for p in ps:
            spa.append(sp_index[i, p])
            spb.append(VEC + p)
            L = max(L, vlev[p] + 1)
        vlev[i] = L
        ftgt.append(VEC + i)
        fst.append(start)
        fcnt.append(len(ps))
        flev.append(L - 1)
    fo2, sfptr = _stage(flev)
    sf_tasks = _pack_tasks(*(np.asarray(x)[fo2] if ftgt else [] for x in
                             (ftgt, fst, fcnt, -np.ones(len(ftgt), dtype=np.int64))))
    # Backward over the sparse rows, after the core: v_i = (v_i - sum_{j>i} U_ij v_j) / U_ii.
    btgt, bst, bcnt, bdiv, blev = [], [], [], [], []
    ulev = np.zeros(n, dtype=np.int64)
    for i in range(c - 1, -1, -1):
        js = [j for j in range(i + 1, n) if F[i, j]]
        start = len(spa)
        L = 0
        for j in js:
            spa.append(sp_index[i, j])
            spb.append(VEC + j)
            L = max(L, ulev[j])          # core entries are final before stage 0
        ulev[i] = L + 1
        btgt.append(VEC + i)
        bst.append(start)
        bcnt.append(len(js))
        bdiv.append(sp_index[i, i])
        blev.append(L)
    bo, sbptr = _stage(blev)
    sb_tasks = _pack_tasks(*(np.asarray(x)[bo] if btgt else [] for x in (btgt, bst, bcnt, bdiv)))
    s_pairs = _pack_pairs(spa, spb)

    arrays = dict(
        sp_src=sp_src, sp_diag=sp_diag, sp_row=sp_row,
        core_src=core_src.ravel(), sch_mask=sch_mask.ravel(), sch_base=sch_base, sch_pairs=sch_pairs,
        core_row=core_row,
        f_tasks=f_tasks, f_ptr=fptr.astype(np.int32), f_pairs=f_pairs,
        sf_tasks=sf_tasks, sf_ptr=sfptr.astype(np.int32),
        sb_tasks=sb_tasks, sb_ptr=sbptr.astype(np.int32),
        s_pairs=s_pairs,
        perm=perm.astype(np.int32),
    )
    return LUPlan(n=n, nnz=nnz, perm=perm, c=c, m=m, tile=tile, F=F, ns=ns, nsch=nsch,
                  nnz_lu=int(F.sum()), arrays=arrays)


# ---------------------------------------------------------------------------
# A CPU model of exactly what the kernels do with a plan: used by the tests to
# check a plan before any kernel runs, never by the product path.
# ---------------------------------------------------------------------------

def _run_tasks(lds, tasks, ptr, pairs):
    tasks = tasks.reshape(-1, 2).astype(np.int64)
    pa = (pairs >> 16).astype(np.int64)
    pb = (pairs & 0xFFFF).astype(np.int64)
    for s in range(len(ptr) - 1):
        for t in range(ptr[s], ptr[s + 1]):
            tg, d = tasks[t, 0] & 0xFFFF, tasks[t, 0] >> 16
            st, cnt = tasks[t, 1] & 0xFFFFF, tasks[t, 1] >> 20
            acc = lds[tg]
            for q in range(st, st + cnt):
                acc -= lds[pa[q]] * lds[pb[q]]
            if d != NO_DIV:
                acc = acc / lds[d]
            lds[tg] = acc


def emulate_factor(plan: LUPlan, A_old: np.ndarray):
    """Factor one dense matrix (original numbering) the way the device does.
    Returns (lds, core): the sparse entries and the m_pad^2 dense core after
    no-pivot LU."""
    c, m, mp = plan.c, plan.m, plan.m_pad
    p = plan.perm
    A = A_old[np.ix_(p, p)]
    ar = plan.arrays
    lds = np.zeros(plan.lds_len)
    si, sj = np.nonzero(plan.F)
    keep = np.minimum(si, sj) < c
    lds[: plan.ns] = A[si[keep], sj[keep]]
    _run_tasks(lds, ar["f_tasks"], ar["f_ptr"], ar["f_pairs"])
    core = np.eye(mp)
    core[:m, :m] = A[c:, c:] * plan.F[c:, c:]
    words = ar["sch_pairs"].astype(np.int64)
    mask = ar["sch_mask"].reshape(NT, 4).astype(np.int64)
    tile = plan.tile
    for t in range(NT):
        sums, acc = [], 0.0
        for q in range(int(ar["sch_base"][t]), int(ar["sch_base"][t + 1])):
            w = words[q]
            acc -= lds[(w >> 16) & 0x7FFF] * lds[w & 0xFFFF]
            if w >> 31:
                sums.append(acc)
                acc = 0.0
        tr, tc = divmod(t, G)
        for s_ in range(tile * tile):
            if (mask[t, s_ >> 5] >> (s_ & 31)) & 1:
                ti, tj = divmod(s_, tile)
                core[ti * G + tr, tj * G + tc] += sums.pop(0)
    for k in range(mp):
        core[k + 1:, k] /= core[k, k]
        core[k + 1:, k + 1:] -= np.outer(core[k + 1:, k], core[k, k + 1:])
    return lds, core


def emulate_solve(plan: LUPlan, lds, core, b_old):
    n, c, mp = plan.n, plan.c, plan.m_pad
    p = plan.perm
    lds = lds.copy()
    V = plan.lds_vec
    lds[V: V + n] = b_old[p]
    lds[V + n:] = 0.0
    ar = plan.arrays
    _run_tasks(lds, ar["sf_tasks"], ar["sf_ptr"], ar["s_pairs"])
    y = lds[V + c: V + c + mp].copy()
    y = np.linalg.solve(np.tril(core, -1) + np.eye(mp), y)
    y = np.linalg.solve(np.triu(core), y)
    lds[V + c: V + c + mp] = y
    _run_tasks(lds, ar["sb_tasks"], ar["sb_ptr"], ar["s_pairs"])
    x = np.empty(n)
    x[p] = lds[V: V + n]
    return x
