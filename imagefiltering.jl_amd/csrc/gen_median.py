#!/usr/bin/env python3
"""Generate median_nets.h: straight-line min/max selection networks for the
mapwindow median kernel (k x k windows, tiles of tx x ty outputs per thread).

median_fast! (src/mapwindow.jl:123-131) returns the ((L+1)÷2)-th smallest of
the L = k*k window values.  For a tile of outputs the network
  1. sorts, per input column, the rows shared by every output row of the
     tile (the column "core"), then merges each output row's extra rows in;
  2. per output, selects the target rank from its k sorted columns with
     forgetful selection: before each Batcher merge, every sorted list is
     cut to the positions that can still hold the target rank, and dropped
     low positions lower the rank;
  3. shares identical comparators between outputs (hash-consing), folds
     comparisons with +inf padding and removes operations no output needs.
Every generated network is checked against sorting on random inputs
(including ties) before it is written.

Usage: python gen_median.py > median_nets.h
"""
import random
import sys

INF = "INF"


class Net:
    def __init__(self, ninputs):
        self.nodes = []          # (op, a, b)
        self.memo = {}
        self.inputs = [("in", i) for i in range(ninputs)]

    def _node(self, op, a, b):
        if a == INF:
            return b if op == "min" else INF
        if b == INF:
            return a if op == "min" else INF
        if a == b:
            return a
        key = (op,) + tuple(sorted((a, b), key=repr))
        if key not in self.memo:
            self.memo[key] = ("n", len(self.nodes))
            self.nodes.append(key)
        return self.memo[key]

    def cx(self, a, b):
        return self._node("min", a, b), self._node("max", a, b)


def oddeven_merge(net, a, b):
    """Batcher odd-even merge of two sorted lists, padded with +inf to a power of 2."""
    n = 1
    while n < max(len(a), len(b)):
        n *= 2
    seq = list(a) + [INF] * (n - len(a)) + list(b) + [INF] * (n - len(b))
    _oem_rec(net, seq, 0, 2 * n, 1)
    return seq


def _oem_rec(net, seq, lo, n, r):
    step = r * 2
    if step < n:
        _oem_rec(net, seq, lo, n, step)
        _oem_rec(net, seq, lo + r, n, step)
        for i in range(lo + r, lo + n - r, step):
            seq[i], seq[i + r] = net.cx(seq[i], seq[i + r])
    else:
        seq[lo], seq[lo + r] = net.cx(seq[lo], seq[lo + r])


def merge(net, a, b):
    if not a:
        return list(b)
    if not b:
        return list(a)
    return oddeven_merge(net, a, b)[: len(a) + len(b)]


def sort(net, xs):
    xs = list(xs)
    if len(xs) <= 1:
        return xs
    h = len(xs) // 2
    return merge(net, sort(net, xs[:h]), sort(net, xs[h:]))


def select(net, lists, r):
    """Rank-r element (0-based) of the union of sorted lists (forgetful)."""
    lists = [list(l) for l in lists if l]
    while True:
        total = sum(len(l) for l in lists)
        kept, drop = [], 0
        for l in lists:
            others = total - len(l)
            lo = min(max(0, r - others), len(l))
            hi = min(len(l) - 1, r)
            drop += lo
            if hi >= lo:
                kept.append(l[lo: hi + 1])
        r -= drop
        lists = kept
        if len(lists) == 1:
            return lists[0][r]
        lists.sort(key=len)
        lists = [merge(net, lists[0], lists[1])] + lists[2:]


def merge_all(net, lists):
    """Full sort of the union of sorted lists by balanced Batcher merges."""
    lists = [list(l) for l in lists if l]
    while len(lists) > 1:
        lists.sort(key=len)
        lists = [merge(net, lists[0], lists[1])] + lists[2:]
    return lists[0] if lists else []


def build(k, tx, ty, xcore=1):
    cols, rows = k + tx - 1, k + ty - 1
    net = Net(cols * rows)
    X = lambda c, r: net.inputs[c + cols * r]
    core_rows = list(range(ty - 1, k))          # rows shared by all output rows
    col_sorted = {}
    for c in range(cols):
        core = sort(net, [X(c, r) for r in core_rows])
        for oy in range(ty):
            extra = [X(c, r) for r in range(oy, oy + k) if r not in core_rows]
            col_sorted[(c, oy)] = merge(net, core, sort(net, extra)) if extra else core
    L = k * k
    rank = (L + 1) // 2 - 1
    outs = []
    for oy in range(ty):
        shared = list(range(tx - 1, k))         # columns in every window of this row
        core = merge_all(net, [col_sorted[(c, oy)] for c in shared]) if (xcore == 1 and tx > 1) else None
        for ox in range(tx):
            cols_w = [col_sorted[(c, oy)] for c in range(ox, ox + k)]
            if xcore == 2:
                outs.append(tableau_select(net, cols_w, k, rank))
            elif core is not None:
                extra = [col_sorted[(c, oy)] for c in range(ox, ox + k) if c not in shared]
                outs.append(select(net, [core] + extra, rank))
            else:
                outs.append(select(net, cols_w, rank))
    return net, outs


def tableau_select(net, cols_w, k, r):
    """Sort the rows of the column-sorted k x k window (columns stay sorted),
    drop entries whose rank bounds exclude r, then select from the rows."""
    L = k * k
    rows_sorted = [sort(net, [cols_w[c][i] for c in range(k)]) for i in range(k)]
    segs, below = [], 0
    for i in range(k):
        seg = []
        for j in range(k):
            lb = (i + 1) * (j + 1) - 1
            ub = L - (k - i) * (k - j)
            if ub < r:
                below += 1
            elif lb <= r:
                seg.append(rows_sorted[i][j])
        if seg:
            segs.append(seg)
    return select(net, segs, r - below)


def live_ops(net, outs):
    need = set()
    stack = [o for o in outs if o[0] == "n"]
    while stack:
        o = stack.pop()
        if o in need:
            continue
        need.add(o)
        op, a, b = net.nodes[o[1]]
        for x in (a, b):
            if x[0] == "n":
                stack.append(x)
    return sorted(need, key=lambda x: x[1])


def check(net, outs, k, tx, ty, trials=300):
    cols, rows = k + tx - 1, k + ty - 1
    rnd = random.Random(1234 + k * 100 + tx * 10 + ty)
    order = live_ops(net, outs)
    for trial in range(trials):
        hi = 4 if trial % 3 == 0 else 1000
        vals = [rnd.randrange(hi) for _ in range(cols * rows)]
        env = {}
        def ev(x):
            return vals[x[1]] if x[0] == "in" else env[x]
        for o in order:
            op, a, b = net.nodes[o[1]]
            env[o] = min(ev(a), ev(b)) if op == "min" else max(ev(a), ev(b))
        i = 0
        for oy in range(ty):
            for ox in range(tx):
                win = sorted(vals[c + cols * r] for r in range(oy, oy + k) for c in range(ox, ox + k))
                want = win[(k * k + 1) // 2 - 1]
                assert ev(outs[i]) == want, (k, tx, ty, trial)
                i += 1


def emit(net, outs, k, tx, ty):
    cols, rows = k + tx - 1, k + ty - 1
    order = live_ops(net, outs)
    name = {}
    lines = []
    for j, o in enumerate(order):
        op, a, b = net.nodes[o[1]]
        def nm(x):
            return "v[%d]" % x[1] if x[0] == "in" else name[x]
        name[o] = "t%d" % j
        lines.append("  const T t%d = %s(%s, %s);" % (j, "kmin" if op == "min" else "kmax", nm(a), nm(b)))
    body = "\n".join(lines)
    rets = "\n".join("  out[%d] = %s;" % (i, ("v[%d]" % o[1]) if o[0] == "in" else name[o])
                     for i, o in enumerate(outs))
    return ("// k=%d tile %dx%d: %d inputs, %d min/max ops, %.1f per output\n"
            "template <typename T> __device__ __forceinline__ void median_net_%d_%d_%d(const T (&v)[%d], T (&out)[%d]) {\n"
            "%s\n%s\n}\n" % (k, tx, ty, cols * rows, len(order), len(order) / (tx * ty),
                             k, tx, ty, cols * rows, tx * ty, body, rets))


CONFIGS = [(3, 4, 2, 2), (5, 2, 2, 1), (7, 2, 2, 1)]  # k, tx, ty, selection mode

if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--survey":
        for k in (3, 5, 7):
            for tx, ty in [(1, 1), (2, 2), (4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (8, 4)]:
                for xc in (0, 1, 2):
                    net, outs = build(k, tx, ty, xc)
                    n = len(live_ops(net, outs))
                    print(k, tx, ty, "xcore", xc, "inputs", (k + tx - 1) * (k + ty - 1), "ops", n,
                          "per-out %.1f" % (n / (tx * ty)))
        sys.exit(0)
    print("// Generated by gen_median.py -- do not edit.")
    print("#pragma once")
    for k, tx, ty, mode in CONFIGS:
        net, outs = build(k, tx, ty, mode)
        check(net, outs, k, tx, ty)
        print(emit(net, outs, k, tx, ty))
