"""The in-bucket sorting networks of gtls_amd/csrc/fold_sort.h sort every input.

The comparator lists are read from the header (the GTLS_CS lines of sort8 and sort16) and
checked by the 0-1 principle (a comparator network sorts every input iff it sorts every
0/1 input): Batcher's 8-key network, the 16-key one, and the 16-key one pruned to its first
12 inputs (comparators that touch inputs 12-15 dropped: those hold +inf keys, which start
at the top and only move up), which the windowed sort uses for 9-12-key buckets. A network
with one comparator removed must fail, so the check has teeth.
"""
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "gtls_amd", "csrc", "fold_sort.h")


def networks():
    src = open(HEADER).read()
    out = {}
    for name in ("sort8", "sort16"):
        i = src.index("void %s(" % name)
        j = src.index("#undef GTLS_CS", i)
        out[name] = [tuple(map(int, m)) for m in re.findall(r"GTLS_CS\((\d+), (\d+)\)", src[i:j])]
    return out


def sorts_all_01(net, n):
    for bits in range(1 << n):
        v = [(bits >> k) & 1 for k in range(n)]
        for a, b in net:
            if v[a] > v[b]:
                v[a], v[b] = v[b], v[a]
        if any(v[k] > v[k + 1] for k in range(n - 1)):
            return False
    return True


def test_networks_sort_every_input():
    nets = networks()
    assert len(nets["sort8"]) == 19 and len(nets["sort16"]) == 63
    assert all(a < b for net in nets.values() for a, b in net)
    assert sorts_all_01(nets["sort8"], 8)
    assert sorts_all_01(nets["sort16"], 16)
    pruned = [(a, b) for a, b in nets["sort16"] if b < 12]
    assert len(pruned) == 42
    assert sorts_all_01(pruned, 12)
    # a dropped comparator is caught
    assert not sorts_all_01(pruned[:20] + pruned[21:], 12)
    assert not sorts_all_01(nets["sort8"][1:], 8)
