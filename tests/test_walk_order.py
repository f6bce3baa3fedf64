"""The commit rule of the GPU Barnes-Hut walk (treebuild_dev.hip, k_bh_walk),
checked on the CPU against the serial depth-first walk whose order the lists
must keep (the host builder's, after PG5/newton_quad/source.cpp:832-867).

The GPU step examines the children of the top 8 stack entries at once and
keeps the work of the entries up to and including the first one that pushes
children; emissions are ranked entries first, octants 7..0 inside an entry.
This restates that rule lane by lane on random octrees with random opening
decisions and compares the emitted particle and cell sequences, and checks
that every stack write stays below the depth bound 7 * (levels - 1) + 8."""
import random

import pytest


def random_tree(levels, rng, n=4096):
    """Nodes [lo, hi, first child, octant mask, level]; children of a node are
    consecutive, in octant order, as the GPU builder numbers them."""
    nodes = [[0, n, -1, 0, 0]]
    i = 0
    while i < len(nodes):
        lo, hi, _, _, lev = nodes[i]
        if hi - lo > 1 and lev < levels:
            cuts = sorted(rng.sample(range(lo + 1, hi), min(7, hi - lo - 1)))
            bounds = [lo] + cuts + [hi]
            octs = sorted(rng.sample(range(8), len(bounds) - 1))
            first, mask = len(nodes), 0
            for (a, b), o in zip(zip(bounds[:-1], bounds[1:]), octs):
                nodes.append([a, b, -1, 0, lev + 1])
                mask |= 1 << o
            nodes[i][2], nodes[i][3] = first, mask
        i += 1
    return nodes


def children(nodes, k):
    lo, hi, first, mask, lev = nodes[k]
    out, r = {}, 0
    for o in range(8):
        if mask >> o & 1:
            out[o] = first + r
            r += 1
    return out


def serial_walk(nodes, nleaf, accept):
    pl, cl, st = [], [], [0]
    while st:
        k = st.pop()
        lo, hi = nodes[k][0], nodes[k][1]
        if hi - lo <= nleaf:
            pl += range(lo, hi)
            continue
        ch = children(nodes, k)
        for o in range(7, -1, -1):
            if o not in ch:
                continue
            c = ch[o]
            if nodes[c][1] - nodes[c][0] == 1:
                pl.append(nodes[c][0])
            elif accept(c):
                cl.append(c)
            else:
                st.append(c)
    return pl, cl


def popc(x):
    return bin(x).count("1")


def wave_walk(nodes, nleaf, accept, depth):
    def entry(k):
        lo, hi, first, mask, lev = nodes[k]
        return (lo, hi, -1) if hi - lo <= nleaf else (first, mask, lev)

    full = (1 << 64) - 1
    stk = [None] * depth
    stk[0], sp = entry(0), 1
    pl, cl, np_, nc = {}, {}, 0, 0
    while sp > 0:
        npop = min(sp, 8)
        lanes = []
        for lane in range(64):
            e, o = lane >> 3, lane & 7
            ent = stk[sp - 1 - e] if e < npop else (0, 0, 0)
            leaf = e < npop and ent[2] < 0
            kind, c = 0, -1
            if e < npop and not leaf and ent[1] >> o & 1:
                c = ent[0] + popc(ent[1] & ((1 << o) - 1))
                kind = 1 if nodes[c][1] - nodes[c][0] == 1 else (2 if accept(c) else 3)
            lanes.append((e, o, ent, leaf, kind, c))
        ms = sum(1 << l for l in range(64) if lanes[l][4] == 3)
        ncommit = (((ms & -ms).bit_length() - 1) >> 3) + 1 if ms else npop
        mp = sum(1 << l for l in range(64) if lanes[l][0] < ncommit and lanes[l][4] == 1)
        mc = sum(1 << l for l in range(64) if lanes[l][0] < ncommit and lanes[l][4] == 2)
        leafcnt = [lanes[8 * q][2][1] - lanes[8 * q][2][0]
                   if lanes[8 * q][0] < ncommit and lanes[8 * q][3] else 0 for q in range(8)]
        mpush = ms & (0xFF << (8 * (ncommit - 1)))
        writes = []
        for lane, (e, o, ent, leaf, kind, c) in enumerate(lanes):
            if e >= ncommit:
                continue
            mine, before = 0xFF << (8 * e), (1 << (8 * e)) - 1
            above = ~((2 << lane) - 1) & full
            pbase = np_ + sum(leafcnt[:e]) + popc(mp & before)
            if kind == 1:
                pos = pbase + popc(mp & mine & above)
                assert pos not in pl
                pl[pos] = nodes[c][0]
            if leaf:
                for p in range(ent[0] + o, ent[1], 8):
                    assert pbase + p - ent[0] not in pl
                    pl[pbase + p - ent[0]] = p
            if kind == 2:
                pos = nc + popc(mc & (before | (mine & above)))
                assert pos not in cl
                cl[pos] = c
            if kind == 3:
                writes.append((sp - ncommit + popc(mpush & above), c))
        sp -= ncommit
        for pos, c in writes:
            assert 0 <= pos < depth
            stk[pos] = entry(c)
        np_ += popc(mp) + sum(leafcnt)
        nc += popc(mc)
        sp += popc(mpush)
    assert sorted(pl) == list(range(len(pl))) and sorted(cl) == list(range(len(cl)))
    return [pl[i] for i in range(len(pl))], [cl[i] for i in range(len(cl))]


@pytest.mark.parametrize("nleaf", [1, 2, 8, 20])
def test_wave_walk_keeps_the_serial_order(nleaf):
    rng = random.Random(nleaf)
    for t in range(40):
        levels = 6 + t % 4
        nodes = random_tree(levels, rng)
        salt = rng.random()
        decide = {}

        def accept(c):
            if c not in decide:
                decide[c] = random.Random(c + salt).random() < 0.55
            return decide[c]

        want = serial_walk(nodes, nleaf, accept)
        got = wave_walk(nodes, nleaf, accept, 7 * (levels - 1) + 8)
        assert got == want, (t, nleaf)
